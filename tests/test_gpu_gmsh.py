"""The step on unstructured Gmsh meshes from the reference's meshes/ directory
(read by the Gmsh 4.1 reader): the same bars as on the Kuhn boxes, with
steps of about a sixth of an edge so that segments cross faces and edges
and particles reach the walls within a few tens of steps."""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import pyoracle as O
from start_amd import gmsh, meshgen
from start_amd.species import EV_J, MASS, gas, species
from start_amd.workloads import Workload

from ._parity import Pair, log

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(__file__), "golden", "gmsh")


def gmsh_workload(name, n=4000, seed=17):
    return mesh_workload(gmsh.read(os.path.join(G, name + ".msh")).mesh, name, n, seed)


def mesh_workload(mesh, name, n=4000, seed=17):
    """Ar+ uniform in the mesh, steps of about a sixth of a typical edge, a few
    % collision chance per step, a linear potential along z."""
    om = O.Mesh(mesh.nodes, mesh.tets, mesh.tris)
    rng = np.random.default_rng(seed)
    lo, hi = mesh.nodes.min(axis=0), mesh.nodes.max(axis=0)
    pts = []
    while sum(len(p) for p in pts) < n:
        p = rng.uniform(lo, hi, size=(4 * n, 3))
        t = om.locate_many(p[:, 0], p[:, 1], p[:, 2])
        pts.append(p[t >= 0])
    pos = np.concatenate(pts)[:n]
    p = mesh.nodes[mesh.tets]
    vol = np.abs(np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))) / 6
    h = np.cbrt(6.0 * vol.mean())  # typical edge length
    ar = species("Ar", "Ar", charge_units=1.0, diameter=4.17e-10)
    dt = 1e-6
    sig = 0.15 * h / dt / np.sqrt(3.0)  # mean step ~ a sixth of an edge
    vel = rng.normal(0.0, sig, size=(n, 3))
    g = gas("Ar", 0.0, 300.0, "VSS")
    g.concentration = 0.5 / (np.pi * ar.diameter ** 2 * sig * dt * 20)  # a few % collision chance per step
    phi = meshgen.linear_potential(mesh, 0.0, 0.05 * MASS["Ar"] * sig ** 2 / EV_J)
    return Workload(name, mesh, [ar], g, dt, 100, seed, pos, vel, np.zeros(n, dtype=np.int32), phi)


@pytest.mark.parametrize("name", ["cube", "two_cubes", "sphere", "plates"])
def test_fp64_parity_on_gmsh_meshes(name):
    """Owners, statuses, hits, triangle counts, x and v bit-exact, node
    charge within 1e-12, over 3 x 20 fused steps."""
    w = gmsh_workload(name)
    pr = Pair(w, steps_per_launch=10)
    first = None
    for k in range(3):
        pr.advance(20)
        r = pr.compare(f"gmsh {name} step {pr.step}")
        first = first or r
    log(f"gmsh {name}: {w.mesh.n_tets} tets, at step 20 owner changed {first['owner_changed']}, "
        f"settled {first['settled']}; at step 60 settled {r['settled']}, deferred {r['deferred']}")
    # segments crossed faces: owners changed, or the particle reached a wall
    assert first["owner_changed"] > 200 and first["owner_changed"] + first["settled"] > 2000
    pr.close()


@pytest.mark.parametrize("name", ["sphere", "plates"])
def test_mixed_on_gmsh_meshes(name):
    """Mixed precision on the unstructured meshes against the FP64 oracle, to
    the mixed bars."""
    w = gmsh_workload(name)
    pr = Pair(w, steps_per_launch=10, precision="mixed")
    pr.advance(40)
    pr.compare_mixed(f"gmsh {name}")
    pr.close()
