"""The C++ step-loop entry points (ParticleDynamicsProcessor::process and
NodeChargeDensityProcessor::gather over the C ABI), driven like the reference's
main loop by shim_check, against the oracle on the same inputs and variates."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as O
from start_amd import species as S
from start_amd import workloads

from ._parity import REL_Q, REL_XV, log
from .test_gpu_movements import check_hit

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TYPES = ["O2", "Ar", "Ne", "He", "Ti", "Al", "Sn", "W", "Au", "Cu", "Ni", "Ag"]


def _run_shim(tmp, w, steps, density, charge_units, diameter):
    n = len(w.pos)
    tmp.mkdir(exist_ok=True)
    w.mesh.nodes.astype(np.float64).tofile(tmp / "nodes.f64")
    w.mesh.tets.astype(np.int32).tofile(tmp / "tets.i32")
    w.mesh.tris.astype(np.int32).tofile(tmp / "tris.i32")
    w.phi.astype(np.float64).tofile(tmp / "phi.f64")
    w.pos.astype(np.float64).tofile(tmp / "pos.f64")
    w.vel.astype(np.float64).tofile(tmp / "vel.f64")
    np.full(n, TYPES.index("Ti"), dtype=np.int32).tofile(tmp / "type.i32")
    np.arange(1000, 1000 + n, dtype=np.uint64).tofile(tmp / "id.u64")
    cfg = {"dt": repr(w.dt), "steps": steps, "model": "hs", "gas": "Ar", "pressure": 0.5,
           "temperature": 300.0, "seed": w.seed, "precision": "fp64", "density": density,
           "diameter_Ti": repr(diameter), "charge_Ti": charge_units}
    (tmp / "config.txt").write_text("".join(f"{k} {v}\n" for k, v in cfg.items()))
    log(f"shim_check: {n} particles, {steps} steps, density {density}")
    r = subprocess.run([os.path.join(ROOT, "start_amd", "shim_check"), str(tmp)], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and "ERROR" not in r.stderr, r.stdout + r.stderr
    return {
        "pos": np.fromfile(tmp / "out_pos.f64").reshape(-1, 3),
        "vel": np.fromfile(tmp / "out_vel.f64").reshape(-1, 3),
        "settled": np.fromfile(tmp / "out_settled.u64", dtype=np.uint64),
        "tri": np.fromfile(tmp / "out_tri_counts.u64", dtype=np.uint64),
        "rho": np.fromfile(tmp / "out_density.f64"),
        "written": np.fromfile(tmp / "out_density_written.i32", dtype=np.int32).astype(bool),
        "summary": (tmp / "out_summary.txt").read_text(),
        "mv_ids": np.fromfile(tmp / "out_mv_ids.u64", dtype=np.uint64),
        "mv_counts": np.fromfile(tmp / "out_mv_counts.i32", dtype=np.int32),
        "mv_xyz": np.fromfile(tmp / "out_mv_xyz.f64").reshape(-1, 3),
    }


@pytest.mark.parametrize("density", ["barycentric", "tet_average"])
def test_cpp_step_loop_matches_oracle(tmp_path, density):
    charge_units, diameter, steps = 2.0, 2.0e-10, 60
    w = workloads.small_box(n_particles=2500, charge_units=charge_units)
    out = _run_shim(tmp_path / density, w, steps, density, charge_units, diameter)
    # the oracle with one species per particle type, as the C++ side sets them up
    om = O.Mesh(w.mesh.nodes, w.mesh.tets, w.mesh.tris)
    E = om.tet_field(w.phi)
    gas_mass = S.MASS["Ar"]
    ks = []
    for t in TYPES:
        q = charge_units * S.E_CHARGE if t == "Ti" else S.CHARGE_UNITS.get(t, 0) * S.E_CHARGE
        d = diameter if t == "Ti" else S.RADIUS[t] + S.RADIUS["Ar"]
        ks.append(O.species_constants(S.MASS[t], q, d, gas_mass, S.VTI["Ar"], S.VSS_ALPHA["Ar"], O.MODEL["HS"]))
    n = len(w.pos)
    st = O.State.new(w.pos, w.vel, np.full(n, TYPES.index("Ti")), np.arange(n, dtype=np.uint64),
                     om.locate_many(w.pos[:, 0], w.pos[:, 1], w.pos[:, 2]))
    sim = O.Sim(om, ks, O.MODEL["HS"], S.concentration(0.5, 300.0), w.dt, w.seed, E=E)
    q = np.zeros(w.mesh.n_nodes)
    moved = np.zeros(n, dtype=np.uint8)
    expect = None  # particleMovementMap as the reference fills it (all 2,500 are tracked)
    for k in range(steps):
        q[:] = 0.0
        x0 = np.stack([st.x, st.y, st.z], 1)
        settled0 = st.status == O.SETTLED
        sim.step(st, k, q, nthreads=8, moved=moved)
        if expect is None:
            expect = {int(i): [] for i in np.nonzero(moved)[0]}
        for i in np.nonzero(moved)[0]:
            expect[int(i)].append(("start", x0[i]))
            if not settled0[i] and st.status[i] == O.SETTLED:
                expect[int(i)].append(("hit", x0[i], np.array([st.x[i], st.y[i], st.z[i]]), int(st.hit[i])))
    assert out["summary"].startswith(f"steps {steps}")
    # the movement map of process()
    assert np.array_equal(out["mv_ids"], np.array(sorted(expect), dtype=np.uint64) + 1000)
    assert np.array_equal(out["mv_counts"], [len(expect[i]) for i in sorted(expect)])
    offs = np.r_[0, np.cumsum(out["mv_counts"])]
    for j, i in enumerate(sorted(expect)):
        for k, e in enumerate(expect[i]):
            p = out["mv_xyz"][offs[j] + k]
            if e[0] == "start":
                assert np.abs(p - e[1]).max() <= REL_XV * np.abs(e[1]).max(), (i, k)
            else:
                check_hit(w.mesh, p, e[1], e[2], e[3], 1e-9)
    xo = np.stack([st.x, st.y, st.z], 1)
    vo = np.stack([st.vx, st.vy, st.vz], 1)
    ex = np.abs(out["pos"] - xo).max(1) / np.linalg.norm(xo, axis=1)
    ev = np.abs(out["vel"] - vo).max(1) / np.maximum(np.linalg.norm(vo, axis=1), 1e-300)
    assert ex.max() <= REL_XV and ev.max() <= REL_XV, (ex.max(), ev.max())
    settled_o = np.sort(np.nonzero(st.status == O.SETTLED)[0] + 1000).astype(np.uint64)
    assert np.array_equal(out["settled"], settled_o)
    assert np.array_equal(out["tri"], sim.tri_count[: w.mesh.n_tris])
    vn = om.node_volumes()
    floor = S.E_CHARGE / vn
    if density == "barycentric":
        ref, wr = q / vn, np.ones(w.mesh.n_nodes, dtype=bool)
    else:
        ref, wr = sim.node_density(st)
    assert np.array_equal(out["written"], wr)
    err = np.abs(out["rho"] - ref) / np.maximum(np.abs(ref), floor)
    assert err.max() <= REL_Q, err.max()
    log(f"shim {density}: ok, settled {len(settled_o)}, max rel x {ex.max():.2e} v {ev.max():.2e} "
        f"rho {err.max():.2e}")
