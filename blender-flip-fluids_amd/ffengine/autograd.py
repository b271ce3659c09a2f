"""torch.autograd wrappers over the engine's differentiable stages.

PressureProjection: the pressure projection of one substep (assembly, MIC(0)
PCG solve, velocity update; pressuresolver.cpp:52-104, 1047-1205) as a
function of the three MAC face grids, for a fixed liquid SDF and fixed face
weights.  Its forward runs FFAMD_pressure_solve, its backward the
hand-written adjoint FFAMD_projection_vjp (the same symmetric Poisson solve).
P2GTransfer and SampleVelocity: the particle-to-grid transfer and the
grid-to-particle interpolation, differentiable in the particle velocities
or face grids and in the positions, backed by FFAMD_p2g_vjp,
FFAMD_sample_velocity_vjp and FFAMD_sample_velocity_position_vjp.
Extrapolate: the layered velocity extrapolation for fixed validity flags
(FFAMD_extrapolate_vjp).  AdvectRK3: the RK3 position update through the
MAC field, backed by FFAMD_advect_rk3_vjp.  RolloutDensity: a run of
substeps from an initial MAC field to the particle density (FFAMD_rollout_*,
with substep checkpoints and recomputation inside the library).
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


def _host(x):
    """float32 host array of a tensor or array"""
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(x, dtype=np.float32)


class P2GTransfer(torch.autograd.Function):
    """Particle velocities (P, 3) and positions (P, 3) -> the three MAC face
    grids of the FLIP transfer (velocityadvector.cpp:467-536).  pos may be a
    tensor that requires grad; the position derivative is that of the
    transfer weights (continuous: the kernel's slope is 0 at its radius)."""

    @staticmethod
    def forward(ctx, vel, pos, n3, dx, radius):
        p, v = _host(pos), _host(vel)
        (u, gv, w), _ = kernels.p2g(n3, dx, p, v, radius)
        ctx.static = (p, v, n3, dx, radius)
        return torch.from_numpy(u), torch.from_numpy(gv), torch.from_numpy(w)

    @staticmethod
    def backward(ctx, gu, gv, gw):
        p, v, n3, dx, radius = ctx.static
        grads = [_host(g) for g in (gu, gv, gw)]
        if ctx.needs_input_grad[1]:
            vbar, xbar = kernels.p2g_vjp(n3, dx, p, radius, grads, vel=v)
            return torch.from_numpy(vbar), torch.from_numpy(xbar), None, None, None
        return torch.from_numpy(kernels.p2g_vjp(n3, dx, p, radius, grads)), None, None, None, None


class SampleVelocity(torch.autograd.Function):
    """The three MAC face grids and positions (P, 3) -> trilinear velocity
    samples (P, 3) (macvelocityfield.cpp:519-645), the G2P interpolation.
    pos may be a tensor that requires grad."""

    @staticmethod
    def forward(ctx, u, v, w, pos, dx):
        n3 = (u.shape[2] - 1, u.shape[1], u.shape[0])
        arrays = [_host(t) for t in (u, v, w)]
        p = _host(pos)
        ctx.static = (n3, dx, p, arrays)
        return torch.from_numpy(kernels.sample_velocity(n3, dx, arrays, p))

    @staticmethod
    def backward(ctx, gvel):
        n3, dx, p, arrays = ctx.static
        vb = _host(gvel)
        u, v, w = kernels.sample_velocity_vjp(n3, dx, p, vb)
        pbar = None
        if ctx.needs_input_grad[3]:
            pbar = torch.from_numpy(kernels.sample_velocity_position_vjp(n3, dx, arrays, p, vb))
        return torch.from_numpy(u), torch.from_numpy(v), torch.from_numpy(w), pbar, None


class Extrapolate(torch.autograd.Function):
    """One face grid -> its layered extrapolation (gridutils.h:95-160) for
    fixed validity flags (same shape, uint8) and layer count."""

    @staticmethod
    def forward(ctx, field, valid, layers):
        ctx.static = (np.ascontiguousarray(valid, dtype=np.uint8), int(layers))
        return torch.from_numpy(kernels.extrapolate(_host(field), valid, layers))

    @staticmethod
    def backward(ctx, gout):
        valid, layers = ctx.static
        return torch.from_numpy(kernels.extrapolate_vjp(_host(gout), valid, layers)), None, None


class AdvectRK3(torch.autograd.Function):
    """Positions (P, 3) and the three MAC face grids -> positions after the
    RK3 advection step (fluidsimulation.cpp:7616-7623), without the collision
    step; differentiable in both the positions and the grids."""

    @staticmethod
    def forward(ctx, pos, u, v, w, dx, dt):
        n3 = (u.shape[2] - 1, u.shape[1], u.shape[0])
        arrays = [t.detach().cpu().numpy().astype(np.float32) for t in (u, v, w)]
        p = pos.detach().cpu().numpy().astype(np.float32)
        ctx.static = (n3, dx, dt, arrays, p)
        return torch.from_numpy(kernels.advect_rk3(n3, dx, arrays, p, dt))

    @staticmethod
    def backward(ctx, gpos):
        n3, dx, dt, arrays, p = ctx.static
        pbar, (gu, gv, gw) = kernels.advect_rk3_vjp(n3, dx, arrays, p, dt,
                                                    gpos.detach().cpu().numpy().astype(np.float32))
        return (torch.from_numpy(pbar), torch.from_numpy(gu), torch.from_numpy(gv), torch.from_numpy(gw), None,
                None)


class RolloutDensity(torch.autograd.Function):
    """Initial MAC face grids (u, v, w) -> the particle density after a
    rollout's substeps, (n_k, n_j, n_i).  `rollout` is a ffengine.rollout.Rollout;
    the forward records the masks the backward holds fixed."""

    @staticmethod
    def forward(ctx, u, v, w, rollout):
        ctx.rollout = rollout
        rho = torch.from_numpy(rollout.forward(_host(u), _host(v), _host(w)))
        ctx.generation = rollout.generation
        return rho

    @staticmethod
    def backward(ctx, grho):
        if ctx.rollout.generation != ctx.generation:
            raise RuntimeError("RolloutDensity: the rollout ran another forward since this one; "
                               "its backward needs the state of the forward it differentiates")
        gu, gv, gw = ctx.rollout.backward(_host(grho))
        return torch.from_numpy(gu), torch.from_numpy(gv), torch.from_numpy(gw), None
