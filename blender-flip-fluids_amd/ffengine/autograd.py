"""torch.autograd wrappers over the engine's differentiable stages.

PressureProjection: the pressure projection of one substep (assembly, MIC(0)
PCG solve, velocity update; pressuresolver.cpp:52-104, 1047-1205) as a
function of the three MAC face grids, for a fixed liquid SDF and fixed face
weights.  Its forward runs FFAMD_pressure_solve, its backward the
hand-written adjoint FFAMD_projection_vjp (the same symmetric Poisson solve).
Tensors live on the host: the entry points take and return host arrays, and
the device work happens inside the library calls.
"""
import numpy as np
import torch

from . import kernels


class PressureProjection(torch.autograd.Function):
    """(u, v, w) -> projected (u, v, w).  Face grids are float32 tensors
    shaped (k, j, i): u (n_k, n_j, n_i + 1), v (n_k, n_j + 1, n_i),
    w (n_k + 1, n_j, n_i).  weights = the three face-weight grids (same
    shapes), liquid_phi = the cell SDF (n_k, n_j, n_i)."""

    @staticmethod
    def forward(ctx, u, v, w, dx, dt, weights, liquid_phi, tol=1e-9):
        n3 = (u.shape[2] - 1, u.shape[1], u.shape[0])
        arrays = [t.detach().cpu().numpy().astype(np.float32) for t in (u, v, w)]
        _, (pu, pv, pw), _, _, _, ok = kernels.pressure_solve(n3, dx, dt, arrays, weights, liquid_phi, tol=tol,
                                                              mode=0)
        if not ok:
            raise RuntimeError("PressureProjection: the pressure solve did not converge")
        ctx.static = (n3, dx, dt, weights, liquid_phi, tol)
        return tuple(torch.from_numpy(np.ascontiguousarray(a)) for a in (pu, pv, pw))

    @staticmethod
    def backward(ctx, gu, gv, gw):
        n3, dx, dt, weights, phi, tol = ctx.static
        grads = [g.detach().cpu().numpy().astype(np.float32) for g in (gu, gv, gw)]
        (ou, ov, ow), _, ok = kernels.projection_vjp(n3, dx, dt, weights, phi, grads, tol=tol)
        if not ok:
            raise RuntimeError("PressureProjection: the adjoint solve did not converge")
        return (torch.from_numpy(ou), torch.from_numpy(ov), torch.from_numpy(ow), None, None, None, None, None)


def pressure_projection(u, v, w, dx, dt, weights, liquid_phi, tol=1e-9):
    return PressureProjection.apply(u, v, w, dx, dt, weights, liquid_phi, tol)
