#!/usr/bin/env python3
"""Throughput of the per-timestep particle kernel on one MI355X.

Workload (large_box): 0.5 m cube, 50^3 hexes x 6 Kuhn tets = 750,000 tets,
132,651 nodes, 30,000 boundary triangles; 1,000,000 Ar+ (+1e), 1 eV
Maxwellian, uniform positions; VHS collisions with Ar at 1 Pa / 300 K
(d_ref 4.17e-10 m, omega 0.81f); E from a 0 / 500 V linear potential;
barycentric charge scatter every step; dt 1e-9 s.  Synthetic, seeded inputs.

A step is one pass of the hot path over all particles: push, move, collision,
relocation, surface hit/settle, charge scatter.  The timed region holds exactly
--steps steps with the particles resident in HBM, bracketed by device
synchronisation.  value = particle-steps processed (not settled, not frozen)
per second.

Prints one JSON line.  --gpus N > 1 is refused: this build runs on one GPU.
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

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
# algorithmic bytes per particle-step: 2*(6s+8) + (3 s_E + 12 s_G + 32) + 32
BYTES_PER_PS = {"fp64": 296, "mixed": 188}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


def gpu_run(w, mode: str, steps: int, warmup: int):
    from start_amd.engine import ParticleInCellStep
    t_setup = time.time()
    eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed, precision=mode,
                             charge_every_step=True)
    eng.load_particles(w.pos, w.vel, w.sp_idx)
    eng.set_node_potential(w.phi)
    eng.synchronize()
    log(f"[bench] {mode}: device set-up {time.time() - t_setup:.1f}s")
    eng.process(warmup)
    eng.synchronize()
    s0 = eng.stats()
    eng.timing(True)
    eng.synchronize()
    t0 = time.perf_counter()
    eng.process(steps)
    eng.synchronize()
    t1 = time.perf_counter()
    kern_ms, launches = eng.timing_read()
    s1 = eng.stats()
    eng.close()
    processed = s1["processed"] - s0["processed"]
    wall = t1 - t0
    return {
        "processed": processed, "wall_s": wall, "ps_per_s": processed / wall,
        "ms_per_step": 1e3 * wall / max(steps, 1),
        "kernel_ms_avg": kern_ms / max(launches, 1), "launches": launches,
        "deferred": s1["deferred"] - s0["deferred"], "settled": s1["settled"],
    }


def cpu_baseline(w, budget_s: float = 12.0, max_steps: int = 200):
    """The oracle (C restatement of the reference step, OpenMP over particles)
    on this host's cores, on the same workload, for a bounded number of steps."""
    from oracle import pyoracle as O
    om = O.Mesh(w.mesh.nodes, w.mesh.tets, w.mesh.tris)
    E = om.tet_field(w.phi)
    g = w.gas
    ks = [O.species_constants(s.mass, s.charge, s.diameter, g.mass, g.omega, g.alpha,
                              O.MODEL[g.model]) for s in w.species]
    tet0 = om.locate_many(w.pos[:, 0], w.pos[:, 1], w.pos[:, 2])
    st = O.State.new(w.pos, w.vel, w.sp_idx, np.arange(len(w.pos), dtype=np.uint64), tet0)
    sim = O.Sim(om, ks, O.MODEL[g.model], g.concentration, w.dt, w.seed, E=E)
    q = np.zeros(w.mesh.n_nodes)
    nth = cpu_threads()
    q[:] = 0
    sim.step(st, 0, q, nthreads=nth)  # warm-up step
    done, processed, t0 = 0, 0, time.perf_counter()
    while done < max_steps and time.perf_counter() - t0 < budget_s:
        q[:] = 0
        processed += sim.step(st, 1 + done, q, nthreads=nth)
        done += 1
    el = time.perf_counter() - t0
    return {"value": processed / el, "unit": "particle-steps/s", "cores": nth, "kind": "port",
            "sample": f"oracle on the same large_box inputs, steps 1..{done} after 1 warm-up "
                      f"step ({processed} particle-steps in {el:.1f} s, OpenMP {nth} threads)"}


def load_traffic(mode: str):
    """HBM bytes per launch of step_main from the committed rocprofv3 summary."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return d.get(mode, {}).get("hbm_bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mode", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the particle step "
              "is measured per GPU and its multi-GPU sharding is described, not run.", file=sys.stderr)
        sys.exit(2)

    from start_amd import workloads
    t = time.time()
    w = workloads.large_box(n_particles=args.particles)
    log(f"[bench] workload built in {time.time() - t:.1f}s: {w.mesh.n_tets} tets, "
        f"{w.mesh.n_nodes} nodes, {w.mesh.n_tris} triangles, {len(w.pos)} particles")

    r = gpu_run(w, args.mode, args.steps, args.warmup)
    log(f"[bench] gpu {args.mode}: {r}")
    per_launch = r["processed"] / max(args.steps, 1)
    B = BYTES_PER_PS[args.mode]
    achieved = B * per_launch / (r["kernel_ms_avg"] * 1e-3) / 1e9 if r["kernel_ms_avg"] > 0 else 0.0
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(args.mode),
            "kernel": "step_main", "bytes_per_particle_step": B,
            "kernel_ms_avg": round(r["kernel_ms_avg"], 5)}

    cpu = None
    if not args.no_cpu:
        t = time.time()
        cpu = cpu_baseline(w)
        log(f"[bench] cpu baseline {cpu} ({time.time() - t:.1f}s)")

    out = {
        "metric": "particle-steps/s on 1 MI355X (FP64 plain vs FP32/batched) + % HBM peak",
        "value": round(r["ps_per_s"], 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.mode == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "large_box: 750,000 tets, 1,000,000 Ar+ (1 eV Maxwellian), VHS Ar 1 Pa "
                               "300 K, E from 0/500 V, charge scatter every step, dt 1e-9 s",
                   "mode": args.mode, "particles": len(w.pos), "tets": w.mesh.n_tets,
                   "particle_steps_timed": r["processed"], "parallelism": "1 GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
