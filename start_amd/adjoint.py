"""Differentiable deposition: the collisionless push and surface-hit map on the
GPU, with its hand-written adjoint behind a torch.autograd.Function.

    model = PlateDeposition(w)              # mesh, particles, field bases
    V = torch.tensor([150.0, 30.0], dtype=torch.float64, requires_grad=True)
    m = DepositionMoments.apply(V, model)   # <x>, <y>, <x^2>, <y^2> on the target plate
    (m[2] - m[0] ** 2).backward()           # V.grad = d(var x)/d(V_top, V_bottom)

V are the two electrode potentials: the field is V[0] E_0 + V[1] E_1 with E_0,
E_1 the fields of unit potential on each electrode.  The moments are over the
particles that settle on the target triangles within the run; their hit points
are x_h = x_a + s (x_b - x_a) on the hit triangle's plane.  The backward pass
is one reverse sweep on the device (start_adjoint_push), which recomputes each
particle's owners backwards in time from its settle record instead of storing
trajectories.  The set of settled particles is held fixed in the derivative
(it changes only at isolated potentials).
"""
from __future__ import annotations

import numpy as np
import torch

from .engine import ParticleInCellStep


class PlateDeposition:
    """One device context holding the mesh, the two field bases and the initial
    particles; every forward() restarts from those particles."""

    def __init__(self, w, precision: str = "fp64", steps: int | None = None, steps_per_launch: int = 100,
                 sort_interval: int = 200):
        self.w = w
        self.steps = int(w.steps if steps is None else steps)
        self.eng = ParticleInCellStep(w.mesh, w.species, w.gas, w.dt, w.seed, precision=precision,
                                      charge_every_step=False, steps_per_launch=steps_per_launch,
                                      sort_interval=sort_interval)
        self.eng.set_field_basis(0, phi=w.phi_top)
        self.eng.set_field_basis(1, phi=w.phi_bottom)
        self.eng.record_settle(True)
        self.target = np.asarray(w.bottom, dtype=bool)
        self.n = len(w.pos)

    def forward(self, V) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        V = np.asarray(V, dtype=np.float64)
        self.eng.load_batch([{"n": self.n, "scale": (float(V[0]), float(V[1])), "concentration": 0.0, "run": 0}],
                            self.w.pos, self.w.vel, self.w.sp_idx)
        self.eng.process(self.steps)
        return self.eng.hits()

    def on_target(self, tri: np.ndarray) -> np.ndarray:
        return (tri >= 0) & self.target[np.clip(tri, 0, None)]

    def backward(self, lam: np.ndarray) -> np.ndarray:
        """dL/dV for lam = dL/d(hit point) (n x 3) of the last forward()."""
        return self.eng.adjoint_push(lam, first_step=0)

    def close(self):
        self.eng.close()


def moments(xh: np.ndarray, sel: np.ndarray) -> np.ndarray:
    x, y = xh[sel, 0], xh[sel, 1]
    return np.array([x.mean(), y.mean(), (x * x).mean(), (y * y).mean()])


def moments_lambda(xh: np.ndarray, sel: np.ndarray, g: np.ndarray) -> np.ndarray:
    """dL/dx_h per particle for L = g . moments (moments over the selected particles)."""
    nb = max(int(sel.sum()), 1)
    lam = np.zeros_like(xh)
    lam[sel, 0] = (g[0] + 2.0 * g[2] * xh[sel, 0]) / nb
    lam[sel, 1] = (g[1] + 2.0 * g[3] * xh[sel, 1]) / nb
    return lam


class DepositionMoments(torch.autograd.Function):
    """V (2 electrode potentials, float64) -> first and second moments of the
    deposited x, y on the target plate; backward by the device adjoint."""

    @staticmethod
    def forward(ctx, V: torch.Tensor, model: PlateDeposition) -> torch.Tensor:
        xh, tri, _ = model.forward(V.detach().cpu().numpy())
        sel = model.on_target(tri)
        ctx.model, ctx.xh, ctx.sel = model, xh, sel
        return torch.as_tensor(moments(xh, sel), dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        lam = moments_lambda(ctx.xh, ctx.sel, grad_out.detach().cpu().numpy().astype(np.float64))
        g = ctx.model.backward(lam)
        return torch.as_tensor(g, dtype=torch.float64), None
