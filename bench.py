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
per second in the headline mode: FP64, one step per launch with the node
charge of every step, the shape of the reference's loop (gather, solve,
process every step), bit-exact with the CPU oracle.  The other modes (mixed
precision, many steps fused per launch) are reported under "modes".

roofline: the streaming per-step kernel is bound by HBM; achieved = the
algorithmic bytes of its launches (296 B per FP64 particle-step, 212 B mixed)
over the median HIP-event duration of >= 5 launches on the library's stream.
Fused launches keep the particle state in registers across steps and are
bound by VALU issue, not HBM: their roofline uses the VALU instructions per
particle-step counted by rocprofv3 (profiles/traffic.json) against the issue
rate of the chip.  traffic = the copy-calibrated HBM bytes per particle-step
of the same kernel and launch shape (profiles/traffic.json) times the
particle-steps of the timed launch.  A fraction above 1 is reported as
invalid and fails the run.

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
# VALU issue rate: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every 4
# cycles (16 lanes) at the 2.4 GHz peak engine clock, in wave-instructions/s
VALU_PEAK = 256 * 4 * 2.4e9 / 4
# algorithmic bytes per particle-step of the streaming step:
#   2*(3s + 3*8 + 8) particle state (positions in s bytes, velocities FP64)
#   + (3 s_E + 12 s_G + 16 + 16) tet record + 4*8 charge
BYTES_PER_PS = {"fp64": 296, "mixed": 212}

MODES = {
    # name: (precision, steps fused per launch, sort interval)
    "fp64_per_step": ("fp64", 1, 50),
    "mixed_per_step": ("mixed", 1, 50),
    "fp64_fused": ("fp64", 200, 200),
    "mixed_fused": ("mixed", 200, 200),
}
HEADLINE = "fp64_per_step"
MIN_TIMED_LAUNCHES = 5


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


def timed_launches(eng, K: int, launches: int):
    """Runs `launches` launches of K steps with HIP events on the library's
    stream; returns (median, mean ms per launch of step_main, particle-steps
    per launch)."""
    s1 = eng.stats()
    eng.timing(True)
    eng.process(K * launches)
    eng.synchronize()
    ms = eng.timing_read_each()
    eng.timing(False)
    s2 = eng.stats()
    return float(np.median(ms)), float(np.mean(ms)), (s2["processed"] - s1["processed"]) / max(len(ms), 1), len(ms)


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
    # duration of the streaming kernel: median over >= 5 launches (one step
    # each, or K fused steps each), events on the library's stream
    n_launch = max(MIN_TIMED_LAUNCHES, 200 if K == 1 else MIN_TIMED_LAUNCHES)
    med_ms, mean_ms, ps_launch, n_timed = timed_launches(eng, K, n_launch)
    s2 = eng.stats()
    eng.close()
    processed = s1["processed"] - s0["processed"]
    wall = t1 - t0
    return {
        "precision": prec, "steps_per_launch": K, "sort_interval": sort,
        "processed": processed, "wall_s": wall, "ps_per_s": processed / wall,
        "ms_per_step": 1e3 * wall / max(steps, 1),
        "launch_ms_median": med_ms, "launch_ms_mean": mean_ms, "launches_timed": n_timed,
        "processed_per_launch": ps_launch, "total_processed": s2["processed"],
        "deferred": s1["deferred"] - s0["deferred"], "settled": s1["settled"],
    }


def accuracy_run(w, steps: int):
    """The mixed path against FP64 on the same inputs and variates, both
    stepped in single-step launches in lockstep (owners compared after every
    step) and compared in full at checkpoints up to `steps`: per-particle
    relative errors of x and v (|d|_inf / |ref|) over the particles whose owner
    and status agree now, and over those that agreed after every step, and the
    count of particles whose owner or status differs."""
    from start_amd.engine import ParticleInCellStep
    engs = {}
    for prec in ("fp64", "mixed"):
        e = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed, precision=prec,
                               charge_every_step=True, steps_per_launch=1, sort_interval=50)
        e.load_particles(w.pos, w.vel, w.sp_idx)
        e.set_node_potential(w.phi)
        engs[prec] = e
    checkpoints = [c for c in (1, 10, 50, 100, 200, 400, 1000, 2000) if c < steps] + [steps]
    rows, done = [], 0
    ever = np.zeros(len(w.pos), dtype=bool)  # owner or status differed after some step
    for cp in checkpoints:
        out = {}
        for _ in range(cp - done):
            own = {}
            for prec, e in engs.items():
                e.process(1)
                own[prec] = e.owners()
            ever |= (own["fp64"][0] != own["mixed"][0]) | (own["fp64"][1] != own["mixed"][1])
        for prec, e in engs.items():
            out[prec] = e.particles()
            out[prec]["charge"] = e.node_charge()
        done = cp
        a, b = out["fp64"], out["mixed"]
        agree = (a["tet"] == b["tet"]) & (a["status"] == b["status"])
        xa = np.stack([a["x"], a["y"], a["z"]], 1)
        xb = np.stack([b["x"], b["y"], b["z"]], 1)
        va = np.stack([a["vx"], a["vy"], a["vz"]], 1)
        vb = np.stack([b["vx"], b["vy"], b["vz"]], 1)
        ex = (np.abs(xa - xb).max(1) / np.linalg.norm(xa, axis=1))[agree]
        # against the extent of the domain: |x| is measured from an arbitrary
        # origin (a corner of the box), so particles near it show large |d|/|x|
        extent = float(np.abs(w.mesh.nodes.max(0) - w.mesh.nodes.min(0)).max())
        ex_dom = (np.abs(xa - xb).max(1) / extent)[agree]
        vn = np.linalg.norm(va, axis=1)
        ev = (np.abs(va - vb).max(1) / np.maximum(vn, 1e-300))[agree]
        # against the thermal speed scale: the post-collision centre-of-mass
        # sum (HSModel.cpp:25-34) nearly cancels for a few particles, whose |v|
        # is then far below the scale of the arithmetic that produced it
        v_scale = float(np.sqrt(np.mean(vn ** 2)))
        ev_scale = (np.abs(va - vb).max(1) / v_scale)[agree]
        qa, qb = a["charge"], b["charge"]
        big = np.abs(qa) > 1e-3 * np.abs(qa).max()
        # the particles whose owner and status agreed after every step so far
        al = ~ever[agree]
        rows.append({
            "steps": cp, "compared": int(agree.sum()), "owner_or_status_differ": int((~agree).sum()),
            "owner_or_status_ever_differed": int(ever.sum()),
            "always_agreed_max_rel_x": float(ex[al].max()),
            "always_agreed_max_rel_v": float(ev[al].max()),
            "always_agreed_max_rel_v_thermal_floor": float(
                (np.abs(va - vb).max(1)[agree][al] / np.maximum(vn[agree][al], v_scale)).max()),
            "max_rel_x": float(ex.max()), "p999_rel_x": float(np.quantile(ex, 0.999)),
            "max_x_err_of_extent": float(ex_dom.max()),
            "max_rel_v": float(ev.max()), "p999_rel_v": float(np.quantile(ev, 0.999)),
            "median_rel_v": float(np.median(ev)),
            "particles_rel_v_above_1e-5": int((ev > 1e-5).sum()),
            "max_rel_v_of_thermal_speed": float(ev_scale.max()),
            "node_charge_max_rel": float((np.abs(qa - qb)[big] / np.abs(qa[big])).max()),
            "settled_fp64": int((a["status"] == 1).sum()), "settled_mixed": int((b["status"] == 1).sum())})
        log(f"[bench] accuracy at {cp} steps: {rows[-1]}")
    for e in engs.values():
        e.close()
    return {"particles": len(w.pos), "thermal_speed_m_s": v_scale, "checkpoints": rows}


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
        # kernel time of step_main, outside the timed region: >= 5 launches
        med_ms, mean_ms, ps_launch, n_timed = timed_launches(eng, K, MIN_TIMED_LAUNCHES)
        eng.close()
        processed = s1["processed"] - s0["processed"]
        out[prec] = {"processed": processed, "wall_s": t1 - t0, "ps_per_s": processed / (t1 - t0),
                     "ms_per_step": 1e3 * (t1 - t0) / steps, "launch_ms_median": med_ms,
                     "launch_ms_mean": mean_ms, "launches_timed": n_timed,
                     "processed_per_launch": ps_launch, "steps_per_launch": K, "precision": prec}
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
        # kernel time of step_main, outside the timed region: 10 launches,
        # each after a field refresh as in the timed loop
        eng.timing(True)
        for k in range(steps, steps + 10 * w.refresh, w.refresh):
            eng.set_node_potential(w.potential(k))
            eng.process(w.refresh)
        eng.synchronize()
        ms = eng.timing_read_each()
        s2 = eng.stats()
        eng.close()
        processed = s1["processed"] - s0["processed"]
        out[prec] = {"processed": processed, "wall_s": t1 - t0, "ps_per_s": processed / (t1 - t0),
                     "ms_per_step": 1e3 * (t1 - t0) / steps, "launch_ms_median": float(np.median(ms)),
                     "launch_ms_mean": float(np.mean(ms)), "launches_timed": len(ms),
                     "processed_per_launch": (s2["processed"] - s1["processed"]) / max(len(ms), 1),
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


def cpp_loop_run(w, steps: int, warmup: int):
    """The drop-in path itself: the C++ ParticleDynamicsProcessor::process and
    NodeChargeDensityProcessor::gather driven by shim_check the way the
    reference's loop calls them (gather, then process, every step;
    ModelingMainDriver.cpp:146-204) on the full-size large box, with the
    density map filled every step.  Once with the default movement map
    (200,000 particles recorded every step, ParticleMovementTracker.hpp:43)
    and once without.  Wall time of `steps` steps after `warmup`."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "start_amd", "shim_check")
    types = ["O2", "Ar", "Ne", "He", "Ti", "Al", "Sn", "W", "Au", "Cu", "Ni", "Ag"]
    out = {}
    with tempfile.TemporaryDirectory(prefix="cpp_loop_", dir=os.environ.get("TMPDIR", "/tmp")) as d:
        w.mesh.nodes.astype(np.float64).tofile(os.path.join(d, "nodes.f64"))
        w.mesh.tets.astype(np.int32).tofile(os.path.join(d, "tets.i32"))
        w.mesh.tris.astype(np.int32).tofile(os.path.join(d, "tris.i32"))
        w.phi.astype(np.float64).tofile(os.path.join(d, "phi.f64"))
        w.pos.astype(np.float64).tofile(os.path.join(d, "pos.f64"))
        w.vel.astype(np.float64).tofile(os.path.join(d, "vel.f64"))
        np.full(len(w.pos), types.index("Ar"), dtype=np.int32).tofile(os.path.join(d, "type.i32"))
        np.arange(len(w.pos), dtype=np.uint64).tofile(os.path.join(d, "id.u64"))
        sp, g = w.species[0], w.gas
        for label, movements in (("no_movement_map", 0), ("movement_map_200k", 200_000)):
            cfg = {"dt": repr(w.dt), "steps": warmup + steps, "model": g.model.lower(), "gas": "Ar",
                   "pressure": 1.0, "temperature": 300.0, "seed": w.seed, "precision": "fp64",
                   "density": "barycentric", "diameter_Ar": repr(sp.diameter),
                   "charge_Ar": repr(sp.charge / 1.602176565e-19), "movements": movements,
                   "timing_warmup": warmup}
            with open(os.path.join(d, "config.txt"), "w") as f:
                f.write("".join(f"{k} {v}\n" for k, v in cfg.items()))
            r = subprocess.run([exe, d], capture_output=True, text=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError("shim_check failed: " + r.stdout + r.stderr)
            t = dict(line.split() for line in open(os.path.join(d, "out_timing.txt")) if line.strip())
            ts, loop = int(t["timed_steps"]), float(t["loop_s"])
            out[label] = {"steps": ts, "ms_per_step": round(1e3 * loop / max(ts, 1), 4),
                          "gather_ms_per_step": round(1e3 * float(t["gather_s"]) / max(ts, 1), 4),
                          "process_ms_per_step": round(1e3 * float(t["process_s"]) / max(ts, 1), 4),
                          "particle_steps_per_s": round(int(t["particle_steps"]) / loop, 1)}
            log(f"[bench] C++ loop {label}: {out[label]}")
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


def roofline(r, key: str):
    """Roofline of the streaming kernel in one run: HBM for single-step
    launches (algorithmic bytes over the median launch duration), VALU issue
    for fused launches (counted VALU instructions per particle-step, from the
    committed counter pass of the same kernel and launch shape).  `key` names
    that counter pass in profiles/traffic.json."""
    prec, K = r["precision"], r["steps_per_launch"]
    dur = r["launch_ms_median"] * 1e-3
    ps = r["processed_per_launch"]
    prof = load_profile(key)
    hbm_ps = prof.get("hbm_bytes_per_particle_step") if prof else None
    common = {"kernel": f"step_main<{'double' if prec == 'fp64' else 'float'}>", "steps_per_launch": K,
              "particle_steps_per_launch": round(ps), "launch_ms": round(r["launch_ms_median"], 4),
              "launch_ms_is": f"median of {r['launches_timed']} launches (HIP events)",
              "traffic": round(hbm_ps * ps) if hbm_ps is not None else None,
              "hbm_bytes_per_particle_step_measured": hbm_ps}
    if K == 1:
        B = BYTES_PER_PS[prec]
        achieved = B * ps / dur / 1e9 if dur > 0 else 0.0
        out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_particle_step": B, **common}
    else:
        valu_ps = prof.get("valu_per_particle_step") if prof else None
        if valu_ps is None:
            return {"bound": "valu", "achieved": None, "peak": VALU_PEAK, "unit": "wave-instructions/s",
                    "frac": None, "note": "no VALU counter pass for this shape in profiles/traffic.json", **common}
        achieved = valu_ps * ps / dur if dur > 0 else 0.0
        out = {"bound": "valu", "achieved": round(achieved, 1), "peak": VALU_PEAK, "unit": "wave-instructions/s",
               "frac": round(achieved / VALU_PEAK, 4), "valu_per_particle_step": valu_ps, **common}
        if key.endswith("_sweep"):
            # the batched kernel's counted issue rate exceeds the one-wave64-
            # instruction-per-4-cycles model (6.8e11 against 6.1e11 per second,
            # transcendental instructions excluded), so that model does not bound
            # it and no fraction is claimed
            out["frac"] = None
            out["note"] = "counted VALU issue above the wave64 4-cycle model: no fraction claimed"
        if hbm_ps is not None and dur > 0:
            out["hbm_frac_measured"] = round(hbm_ps * ps / dur / 1e9 / HBM_PEAK_GBS, 4)
    if out["frac"] is not None and out["frac"] > 1.0:
        out["valid"] = False
        log(f"[bench] INVALID roofline for {key}: frac {out['frac']} > 1")
    return out


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


def load_profile(key: str):
    """Per-particle-step counters of one kernel and launch shape from the
    committed rocprofv3 counter passes (profiles/traffic.json)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(key)
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
    ap.add_argument("--cpp-loop-steps", type=int, default=200,
                    help="steps of the C++ gather/process loop leg (shim_check, full large box); 0 skips it")
    ap.add_argument("--accuracy-steps", type=int, default=400,
                    help="steps of the mixed-vs-FP64 per-result error leg; 0 skips it")
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
    acc = None
    if args.accuracy_steps > 0:
        t = time.time()
        acc = accuracy_run(w, args.accuracy_steps)
        log(f"[bench] accuracy ({time.time() - t:.1f}s): {acc}")
    cpp = None
    if args.cpp_loop_steps > 0:
        cpp = cpp_loop_run(w, args.cpp_loop_steps, 20)
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
    if acc:
        out["mixed_vs_fp64"] = acc
    if cpp:
        out["cpp_loop"] = {"workload": "large_box at full size through ParticleDynamicsProcessor::process and "
                                       "NodeChargeDensityProcessor::gather (shim_check), FP64, density map every step",
                           **cpp}
        if "fp64_per_step" in results:
            out["cpp_loop"]["vs_fp64_per_step"] = {
                k: round(v["particle_steps_per_s"] / results["fp64_per_step"]["ps_per_s"], 3) for k, v in cpp.items()}
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
    bad = []

    def walk(o, path):
        if isinstance(o, dict):
            if o.get("valid") is False:
                bad.append(path)
            for k, v in o.items():
                walk(v, f"{path}.{k}")
    walk(out, "")
    if bad:
        log(f"[bench] roofline fraction above 1 (invalid measurement) at {bad}")
        sys.exit(3)


if __name__ == "__main__":
    main()
