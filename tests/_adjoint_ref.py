"""FP64 reference gradient of the deposition hit points with respect to the two
electrode potentials: forward-mode (tangent) differentiation carried along the
oracle's own discrete map, step by step.

For each particle the tangents dv/dV_j and dx/dV_j (j = top, bottom) follow
    v_{k+1} = v_k + (q/m) E(t_k) dt   ->  dv += (q/m) E_j(t_k) dt   (pushed steps)
    x_{k+1} = x_k + v_{k+1} dt        ->  dx += dv dt
and at the settling step the hit point x_h = x_a + s D on the triangle plane
n.x = n.A (D = x_{k+1} - x_k, s = n.(A - x_a)/n.D) gives
    dx_h = P (dx_a + s dt dv_{k+1}),  P = I - D n^T / (n.D).
The oracle supplies the states, owners and the hit triangle of every step.
"""
from __future__ import annotations

import numpy as np

from oracle import pyoracle as O


def run(w, V, steps=None, nthreads=8):
    om = O.Mesh(w.mesh.nodes, w.mesh.tets, w.mesh.tris)
    E0 = om.tet_field(w.phi_top)
    E1 = om.tet_field(w.phi_bottom)
    E = V[0] * E0 + V[1] * E1
    sp = w.species[0]
    k = O.species_constants(sp.mass, sp.charge, sp.diameter, w.gas.mass, w.gas.omega, w.gas.alpha, 0)
    sim = O.Sim(om, [k], 0, 0.0, w.dt, w.seed, E=E)
    n = len(w.pos)
    st = O.State.new(w.pos, w.vel, w.sp_idx, np.arange(n, dtype=np.uint64),
                     om.locate_many(w.pos[:, 0], w.pos[:, 1], w.pos[:, 2]))
    qm_dt = (sp.charge / sp.mass) * w.dt
    dv = np.zeros((n, 3, 2))
    dx = np.zeros((n, 3, 2))
    xh = np.zeros((n, 3))
    dxh = np.zeros((n, 3, 2))
    hstep = np.full(n, -1, dtype=np.int64)
    tris = w.mesh.tris
    nodes = w.mesh.nodes
    for kstep in range(w.steps if steps is None else steps):
        x0 = np.stack([st.x, st.y, st.z], 1)
        t0 = st.tet.copy()
        s0 = st.status.copy()
        sim.step(st, kstep, nthreads=nthreads)
        x1 = np.stack([st.x, st.y, st.z], 1)
        active = (s0 == O.ACTIVE) & np.any(x1 != x0, axis=1)
        pushed = active & (t0 >= 0)
        tp = t0[pushed]
        dv[pushed, :, 0] += qm_dt * E0[tp]
        dv[pushed, :, 1] += qm_dt * E1[tp]
        now = (s0 == O.ACTIVE) & (st.status == O.SETTLED)
        if now.any():
            xa, D = x0[now], x1[now] - x0[now]
            T = tris[st.hit[now]]
            A, B, C = nodes[T[:, 0]], nodes[T[:, 1]], nodes[T[:, 2]]
            nrm = np.cross(B - A, C - A)
            nD = np.einsum("ij,ij->i", nrm, D)
            s = np.einsum("ij,ij->i", nrm, A - xa) / nD
            xh[now] = xa + s[:, None] * D
            hstep[now] = kstep
            for j in range(2):
                u = dx[now, :, j] + (s * w.dt)[:, None] * dv[now, :, j]
                dxh[now, :, j] = u - D * (np.einsum("ij,ij->i", nrm, u) / nD)[:, None]
        dx[active] += dv[active] * w.dt
    return {"xh": xh, "dxh": dxh, "tri": np.where(st.status == O.SETTLED, st.hit, -1), "step": hstep,
            "state": st}
