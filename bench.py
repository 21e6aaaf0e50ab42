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
per second, in the headline mode (FP64, bit-exact with the CPU oracle); the
other modes are reported under "modes".

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
# algorithmic bytes per particle-step of the streaming step:
#   2*(6s + 8) particle state + (3 s_E + 12 s_G + 16 + 16) tet record + 4*8 charge
BYTES_PER_PS = {"fp64": 296, "mixed": 188}

MODES = {
    # name: (precision, steps fused per launch, sort interval)
    "fp64_fused": ("fp64", 200, 200),
    "fp64_per_step": ("fp64", 1, 50),
    "mixed_fused": ("mixed", 200, 200),
    "mixed_per_step": ("mixed", 1, 50),
}
HEADLINE = "fp64_fused"


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
    prec, K, sort = MODES[mode]
    t_setup = time.time()
    eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed, precision=prec,
                             charge_every_step=True, steps_per_launch=K, sort_interval=sort)
    eng.load_particles(w.pos, w.vel, w.sp_idx)
    eng.set_node_potential(w.phi)
    eng.synchronize()
    log(f"[bench] {mode}: device set-up {time.time() - t_setup:.1f}s")
    eng.process(warmup)
    eng.synchronize()
    s0 = eng.stats()
    # throughput: no event records in the timed region
    t0 = time.perf_counter()
    eng.process(steps)
    eng.synchronize()
    t1 = time.perf_counter()
    s1 = eng.stats()
    # kernel duration of step_main with HIP events on the library's stream
    tsteps = min(steps, max(K, 200 if K == 1 else 2 * K))
    eng.timing(True)
    eng.process(tsteps)
    eng.synchronize()
    kern_ms, launches = eng.timing_read()
    s2 = eng.stats()
    eng.close()
    processed = s1["processed"] - s0["processed"]
    wall = t1 - t0
    return {
        "precision": prec, "steps_per_launch": K, "sort_interval": sort,
        "processed": processed, "wall_s": wall, "ps_per_s": processed / wall,
        "ms_per_step": 1e3 * wall / max(steps, 1),
        "kernel_ms_per_launch": kern_ms / max(launches, 1), "launches_timed": launches,
        "processed_per_launch": (s2["processed"] - s1["processed"]) / max(launches, 1),
        "deferred": s1["deferred"] - s0["deferred"], "settled": s1["settled"],
    }


def sweep_run(steps: int, warmup: int, trial_steps: int = 200):
    """The batched sweep: 64 runs of uneven size on the large box mesh in the
    same launches (mixed precision), launch shape from the autotuner, steps in
    captured hipGraphs; then the same sweep in FP64 for the deposition
    histogram check."""
    from start_amd import workloads
    from start_amd.engine import ParticleInCellStep
    t = time.time()
    sw = workloads.sweep_box()
    n = len(sw.pos)
    log(f"[bench] sweep: {len(sw.runs)} runs, {n} particles, built in {time.time() - t:.1f}s")
    out = {}
    hists = {}
    for prec in ("mixed", "fp64"):
        t = time.time()
        eng = ParticleInCellStep(sw.mesh, sw.species, sw.gas, sw.dt, sw.seed, precision=prec,
                                 steps_per_launch=100, sort_interval=200, graph=True)
        eng.set_field_basis(0, phi=sw.phi_unit)
        eng.load_batch(sw.runs, sw.pos, sw.vel, sw.sp_idx)
        eng.synchronize()
        log(f"[bench] sweep {prec}: device set-up {time.time() - t:.1f}s")
        if prec == "mixed":
            t = time.time()
            cfg, ms = eng.autotune(trial_steps=trial_steps)
            eng.set_launch(**{**cfg, "graph": True})
            log(f"[bench] sweep autotune ({time.time() - t:.1f}s): {cfg}, {ms:.4f} ms/step in trials")
            out["autotuned"] = cfg
        else:
            eng.set_launch(**{**out["autotuned"], "graph": True})
        eng.process(warmup)
        eng.synchronize()
        s0 = eng.stats()
        t0 = time.perf_counter()
        eng.process(steps)
        eng.synchronize()
        t1 = time.perf_counter()
        s1 = eng.stats()
        hists[prec] = eng.tri_counts_runs().astype(np.float64)
        K = eng.get_launch()["steps_per_launch"]
        eng.timing(True)  # kernel time of step_main, outside the timed region
        eng.process(max(K, 2 * K))
        eng.synchronize()
        kern_ms, launches = eng.timing_read()
        s2 = eng.stats()
        eng.close()
        processed = s1["processed"] - s0["processed"]
        out[prec] = {"processed": processed, "wall_s": t1 - t0, "ps_per_s": processed / (t1 - t0),
                     "ms_per_step": 1e3 * (t1 - t0) / steps, "kernel_ms_per_launch": kern_ms / max(launches, 1),
                     "processed_per_launch": (s2["processed"] - s1["processed"]) / max(launches, 1),
                     "steps_per_launch": K}
        log(f"[bench] sweep {prec}: {out[prec]}")
    h64 = hists["fp64"]
    out["deposition_l1_vs_fp64"] = float(np.abs(hists["mixed"] - h64).sum() / max(h64.sum(), 1.0))
    out["deposited"] = int(h64.sum())
    out["runs"] = len(sw.runs)
    out["particles"] = n
    return out


def chamber_run(steps: int):
    """The production-like chamber: 16M Ti neutrals and Ar+, VSS Ar 0.3 Pa,
    charge every step, the node potential handed over from the host every 10
    steps (the launches are 10 steps long, so each refresh lands between two
    launches), substrate hits histogrammed.  Mixed precision is the measured
    configuration; the FP64 run gives the substrate histogram it is checked
    against."""
    from start_amd import workloads
    from start_amd.engine import ParticleInCellStep
    t = time.time()
    w = workloads.chamber(steps=steps)
    sub = w.chamber.substrate
    log(f"[bench] chamber: {w.mesh.n_tets} tets, {w.mesh.n_nodes} nodes, {int(sub.sum())} substrate "
        f"triangles, {len(w.pos)} particles, built in {time.time() - t:.1f}s")
    out, hists = {}, {}
    for prec in ("mixed", "fp64"):
        t = time.time()
        eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed, precision=prec, charge_every_step=True,
                                 steps_per_launch=w.refresh, sort_interval=50, graph=True)
        eng.load_particles(w.pos, w.vel, w.sp_idx)
        eng.set_node_potential(w.potential(0))
        eng.synchronize()
        log(f"[bench] chamber {prec}: device set-up {time.time() - t:.1f}s")
        s0 = eng.stats()
        t0 = time.perf_counter()
        for k in range(0, steps, w.refresh):
            if k:
                eng.set_node_potential(w.potential(k))
            eng.process(min(w.refresh, steps - k))
        eng.synchronize()
        t1 = time.perf_counter()
        s1 = eng.stats()
        hists[prec] = eng.tri_counts()[sub].astype(np.float64)
        eng.timing(True)  # kernel time of step_main, outside the timed region
        for k in range(steps, steps + 10 * w.refresh, w.refresh):
            eng.set_node_potential(w.potential(k))
            eng.process(w.refresh)
        eng.synchronize()
        kern_ms, launches = eng.timing_read()
        s2 = eng.stats()
        eng.close()
        processed = s1["processed"] - s0["processed"]
        out[prec] = {"processed": processed, "wall_s": t1 - t0, "ps_per_s": processed / (t1 - t0),
                     "ms_per_step": 1e3 * (t1 - t0) / steps, "kernel_ms_per_launch": kern_ms / max(launches, 1),
                     "processed_per_launch": (s2["processed"] - s1["processed"]) / max(launches, 1),
                     "steps_per_launch": w.refresh, "settled": s1["settled"], "deferred": s1["deferred"] - s0["deferred"],
                     "precision": prec}
        log(f"[bench] chamber {prec}: {out[prec]}")
    h64 = hists["fp64"]
    out["substrate_deposits"] = int(h64.sum())
    out["substrate_l1_mixed_vs_fp64"] = float(np.abs(hists["mixed"] - h64).sum() / max(h64.sum(), 1.0))
    out["particles"] = len(w.pos)
    out["tets"] = int(w.mesh.n_tets)
    out["nodes"] = int(w.mesh.n_nodes)
    return out


def adjoint_run(reps: int = 3):
    """Forward (collisionless push, surface-hit map) and reverse sweep on the
    parallel-plate workload: 262,144 Cu+, 202,878 tets, 1,000 steps."""
    from start_amd import workloads
    from start_amd.adjoint import PlateDeposition, moments_lambda
    w = workloads.plates()
    model = PlateDeposition(w)
    V = np.array([150.0, 30.0])
    model.forward(V)  # warm-up
    fwd, bwd = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        xh, tri, step = model.forward(V)
        t1 = time.perf_counter()
        sel = model.on_target(tri)
        lam = moments_lambda(xh, sel, np.array([0.0, 0.0, 1.0, 1.0]))
        t2 = time.perf_counter()
        g = model.backward(lam)
        t3 = time.perf_counter()
        fwd.append(t1 - t0)
        bwd.append(t3 - t2)
    stats = model.eng.stats()
    reverse_steps = int((step[sel] + 1).sum())
    model.close()
    f, b = min(fwd), min(bwd)
    return {"workload": "plates: 262,144 Cu+ (lognormal 10 eV, cosine law), 202,878 tets, 1,000 steps of 1e-8 s, "
                        "V = (150, 30) V",
            "forward_s": round(f, 5), "forward_particle_steps_per_s": round(stats["processed"] / f, 1),
            "backward_s": round(b, 5), "backward_particle_steps_per_s": round(reverse_steps / b, 1),
            "on_target": int(sel.sum()), "grad": [float(g[0]), float(g[1])],
            "note": "forward includes the particle upload and owner search; backward includes the "
                    "lambda upload; best of %d" % reps}


def roofline(r, mode):
    prec = MODES[mode][0] if mode in MODES else r["precision"]
    B = BYTES_PER_PS[prec]
    dur = r["kernel_ms_per_launch"] * 1e-3
    achieved = B * r["processed_per_launch"] / dur / 1e9 if dur > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(mode),
            "kernel": f"step_main<{'double' if prec == 'fp64' else 'float'}>",
            "bytes_per_particle_step": B, "steps_per_launch": r["steps_per_launch"],
            "particle_steps_per_launch": round(r["processed_per_launch"]),
            "launch_ms": round(r["kernel_ms_per_launch"], 4)}


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
    sim.step(st, 0, q, nthreads=nth)  # warm-up step
    done, processed, t0 = 0, 0, time.perf_counter()
    while done < max_steps and time.perf_counter() - t0 < budget_s:
        q[:] = 0
        processed += sim.step(st, 1 + done, q, nthreads=nth)
        done += 1
    el = time.perf_counter() - t0
    # the same oracle on one thread, a few more steps of the same state
    one, t1, k1 = 0, time.perf_counter(), 0
    while k1 < 5 and time.perf_counter() - t1 < budget_s / 3:
        q[:] = 0
        one += sim.step(st, 1 + done + k1, q, nthreads=1)
        k1 += 1
    el1 = time.perf_counter() - t1
    return {"value": round(processed / el, 1), "unit": "particle-steps/s", "cores": nth, "kind": "port",
            "sample": f"oracle on the same large_box inputs, steps 1..{done} after 1 warm-up "
                      f"step ({processed} particle-steps in {el:.1f} s, OpenMP {nth} threads)",
            "single_thread": {"value": round(one / el1, 1), "steps": k1, "particle_steps": one},
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def load_traffic(mode: str):
    """HBM bytes per launch of step_main, from the committed rocprofv3 PMC summary."""
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
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--modes", default=",".join(MODES), help="comma-separated subset of " + ",".join(MODES))
    ap.add_argument("--mode-steps", type=int, default=400, help="timed steps of the non-headline modes")
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--sweep-steps", type=int, default=500,
                    help="timed steps of the batched sweep (64 runs); 0 skips it")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--chamber-steps", type=int, default=1000,
                    help="steps of the 16M-particle chamber leg (mixed, then FP64); 0 skips it")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES, 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "copies": 5}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the particle step "
              "is measured per GPU and its multi-GPU sharding is described, not run.", file=sys.stderr)
        sys.exit(2)

    from start_amd import workloads
    t = time.time()
    w = workloads.large_box(n_particles=args.particles)
    log(f"[bench] workload built in {time.time() - t:.1f}s: {w.mesh.n_tets} tets, "
        f"{w.mesh.n_nodes} nodes, {w.mesh.n_tris} triangles, {len(w.pos)} particles")

    modes = [m for m in args.modes.split(",") if m]
    results = {}
    for m in modes:
        steps = args.steps if m == HEADLINE else args.mode_steps
        results[m] = gpu_run(w, m, steps, args.warmup)
        log(f"[bench] {m}: {results[m]}")
    sweep = None
    if args.sweep_steps > 0:
        sweep = sweep_run(args.sweep_steps, args.warmup)
    adj = None
    if not args.no_adjoint:
        adj = adjoint_run()
        log(f"[bench] adjoint: {adj}")
    chamber = None
    if args.chamber_steps > 0:
        chamber = chamber_run(args.chamber_steps)
    if not results:  # a sweep-, adjoint- or chamber-only run (profiling)
        print(json.dumps({"sweep": sweep, "adjoint": adj, "chamber": chamber}), flush=True)
        return
    head = HEADLINE if HEADLINE in results else modes[0]
    r = results[head]

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
        "steps": args.steps if head == HEADLINE else args.mode_steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if MODES[head][0] == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "large_box: 750,000 tets, 1,000,000 Ar+ (1 eV Maxwellian), VHS Ar 1 Pa "
                               "300 K, E from 0/500 V, charge scatter every step, dt 1e-9 s",
                   "mode": head, "steps_per_launch": MODES[head][1], "particles": len(w.pos),
                   "tets": w.mesh.n_tets, "particle_steps_timed": r["processed"], "parallelism": "1 GPU"},
        "roofline": roofline(r, head),
        "cpu_baseline": cpu,
        "modes": {m: {"value": round(v["ps_per_s"], 1), "ms_per_step": round(v["ms_per_step"], 5),
                      "steps": args.steps if m == HEADLINE else args.mode_steps,
                      "precision": v["precision"], "steps_per_launch": v["steps_per_launch"],
                      "roofline": roofline(v, m)} for m, v in results.items()},
    }
    if cpu:
        out["speedup_vs_cpu_baseline"] = round(r["ps_per_s"] / cpu["value"], 1)
    if sweep:
        m, f = sweep["mixed"], sweep["fp64"]
        m["precision"], f["precision"] = "mixed", "fp64"
        out["sweep"] = {
            "workload": "sweep_box: 64 runs on the 750,000-tet box, counts log-uniform in [1e4, 2e6] "
                        f"({sweep['particles']} particles), 0-1000 V, 0.1-5 Pa Ar VSS, dt 1e-9 s",
            "steps": args.sweep_steps, "warmup": args.warmup, "autotuned": sweep["autotuned"],
            "mixed": {"value": round(m["ps_per_s"], 1), "ms_per_step": round(m["ms_per_step"], 5),
                      "roofline": roofline(m, "mixed_sweep")},
            "fp64": {"value": round(f["ps_per_s"], 1), "ms_per_step": round(f["ms_per_step"], 5),
                     "roofline": roofline(f, "fp64_sweep")},
            "deposited": sweep["deposited"],
            "deposition_l1_vs_fp64": sweep["deposition_l1_vs_fp64"],
            "mixed_vs_headline": round(m["ps_per_s"] / r["ps_per_s"], 3),
        }
        if "fp64_per_step" in results:
            out["sweep"]["mixed_vs_fp64_per_step"] = round(m["ps_per_s"] / results["fp64_per_step"]["ps_per_s"], 3)
    if adj:
        out["adjoint"] = adj
    if chamber:
        m, f = chamber["mixed"], chamber["fp64"]
        out["chamber"] = {
            "workload": f"chamber: cylinder r 0.15 m h 0.3 m, {chamber['tets']} tets, {chamber['nodes']} nodes, "
                        f"{chamber['particles']} Ti (Thompson U_b 4.89 eV, cosine) + Ar+ (1 eV), VSS Ar 0.3 Pa, "
                        "charge every step, node potential refreshed every 10 steps, dt 5e-8 s",
            "steps": args.chamber_steps,
            "mixed": {"value": round(m["ps_per_s"], 1), "ms_per_step": round(m["ms_per_step"], 5),
                      "settled": m["settled"], "roofline": roofline(m, "mixed_chamber")},
            "fp64": {"value": round(f["ps_per_s"], 1), "ms_per_step": round(f["ms_per_step"], 5),
                     "settled": f["settled"], "roofline": roofline(f, "fp64_chamber")},
            "substrate_deposits": chamber["substrate_deposits"],
            "substrate_l1_mixed_vs_fp64": chamber["substrate_l1_mixed_vs_fp64"],
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
