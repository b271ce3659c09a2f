"""The differentiable rollout and the adjoints it chains.

Pins:
- the extrapolation's adjoint against the transpose of the oracle's
  Jacobian (one oracle extrapolation per face of a small array);
- the sampling's position cotangent against central differences of a
  float64 trilinear sample, itself pinned to the oracle's float sample;
- one rollout substep against one engine substep on the same particles:
  positions and velocities bit for bit (the rollout runs the engine's
  kernels in the engine's order);
- the rollout gradient does not depend on the checkpoint spacing (the
  recomputation is deterministic), bit for bit;
- the rollout gradient against central differences of the same map
  (forwards with the recorded masks), per face and along a direction;
- the torch.autograd wrapper's gradient equals the entry point's.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from ffengine import kernels as K
from ffengine import scenes
from ffengine.fluidsimulation import FluidSimulation
from ffengine.rollout import Rollout

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _shapes(n3):
    ni, nj, nk = n3
    return [(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]


def test_extrapolate_vjp_is_oracle_jacobian_transposed():
    rng = np.random.default_rng(31)
    shape = (6, 7, 8)  # (wk, wj, wi)
    valid = (rng.random(shape) < 0.15).astype(np.uint8)
    layers = 3
    n = int(np.prod(shape))
    say(f"extrapolation Jacobian: {n} oracle extrapolations")
    J = np.zeros((n, n))
    for f in range(n):
        e = np.zeros(n, np.float32)
        e[f] = 1.0
        J[:, f] = O.extrapolate(e.reshape(shape), valid, layers).astype(np.float64).ravel()
    for trial in range(3):
        b = rng.standard_normal(shape).astype(np.float32)
        want = (J.T @ b.astype(np.float64).ravel()).reshape(shape)
        got = K.extrapolate_vjp(b, valid, layers).astype(np.float64)
        err = np.abs(got - want).max() / np.abs(want).max()
        say(f"  trial {trial}: max |error| / max |J^T b| = {err:.3g}")
        assert err <= 1e-6


def _sample64(n3, dx, uvw, pos):
    """Trilinear MAC sample in float64 (macvelocityfield.cpp:631-645):
    component d at p reads its face grid at p shifted by -dx/2 along the
    other two axes; 0 outside the domain."""
    ni, nj, nk = n3
    pos = pos.astype(np.float64)
    out = np.zeros_like(pos)
    inside = np.all((pos >= 0) & (pos < np.array([ni, nj, nk]) * dx), axis=1)
    for d, a in enumerate(uvw):
        a = a.astype(np.float64)
        wk, wj, wi = a.shape
        s = pos.copy()
        s[:, [e for e in range(3) if e != d]] -= 0.5 * dx
        g = s / dx
        c = np.floor(g).astype(int)
        t = g - c
        acc = np.zeros(len(pos))
        for q in range(8):
            o = np.array([q & 1, (q >> 1) & 1, q >> 2])
            idx = c + o
            ok = np.all((idx >= 0) & (idx < np.array([wi, wj, wk])), axis=1)
            wgt = np.prod(np.where(o == 1, t, 1 - t), axis=1)
            val = np.zeros(len(pos))
            val[ok] = a[idx[ok, 2], idx[ok, 1], idx[ok, 0]]
            acc += wgt * val
        out[:, d] = np.where(inside, acc, 0.0)
    return out


def test_sample_position_vjp_central_differences():
    n = 10
    n3, dx = (n, n, n), 1.0 / n
    rng = np.random.default_rng(32)
    uvw = [rng.standard_normal(s).astype(np.float32) for s in _shapes(n3)]
    pos = (rng.random((400, 3)) * 0.96 + 0.02).astype(np.float32)
    f64 = _sample64(n3, dx, uvw, pos)
    ora = O.sample_velocity(n3, dx, uvw, pos).astype(np.float64)
    pin = np.abs(f64 - ora).max() / np.abs(f64).max()
    say(f"float64 trilinear sample against the oracle: {pin:.3g}")
    assert pin <= 1e-6
    vb = rng.standard_normal(pos.shape).astype(np.float32)
    xbar = K.sample_velocity_position_vjp(n3, dx, uvw, pos, vb).astype(np.float64)
    h = 1e-7
    base = pos.astype(np.float64)
    fd = np.zeros_like(base)
    for e in range(3):
        a, b = base.copy(), base.copy()
        a[:, e] += h
        b[:, e] -= h
        fd[:, e] = ((_sample64(n3, dx, uvw, a) - _sample64(n3, dx, uvw, b)) * vb.astype(np.float64)).sum(1) / (2 * h)
    err = np.abs(fd - xbar).max() / np.abs(xbar).max()
    say(f"sampling position cotangent: max |fd - adjoint| / max |adjoint| = {err:.3g}")
    assert err <= 1e-5


def _ball(n, seed, center=(0.5, 0.55, 0.5), radius=0.22):
    """8 jittered particles per cell in the cells whose centre lies in the
    ball, float32 (P, 3)."""
    pos = scenes.dam_break_particles(n, seed=seed, fraction=(1.0, 1.0, 1.0))
    dx = 1.0 / n
    cell = (np.floor(pos / dx) + 0.5) * dx
    keep = np.linalg.norm(cell - np.array(center), axis=1) < radius
    return np.ascontiguousarray(pos[keep])


@pytest.mark.parametrize("obstacle", [False, True, "friction"])
def test_rollout_substep_matches_engine(obstacle):
    """One substep of the rollout equals the engine's; with a sphere
    MeshObject in the engine and the same solid (FFAMD_solid_phi_meshes, same
    rand() seed) given to the rollout as its nodal solid.  With friction on
    the MeshObject the rollout (a frictionless solid) refuses it."""
    from ffengine import MeshObject, TriangleMesh
    n = 24
    dx = 1.0 / n
    pos = _ball(n, 33)
    init = scenes.noise_mac_field(n, seed=34, amplitude=0.3, correlation=4.0)
    vel = K.sample_velocity((n, n, n), dx, init, pos)
    dt = 1.0 / 120.0
    sphere = scenes.sphere_mesh((0.5, 0.2, 0.5), 0.1, 2)
    sim = FluidSimulation(n, n, n, dx)
    sim.add_body_force(0.0, -9.81, 0.0)
    sim.PICFLIP_ratio = 0.05
    sim.CFL_condition_number = 5
    sim.enable_surface_reconstruction = False
    sim.pressure_solver_tolerance = 1e-9
    if obstacle:
        mo = MeshObject(n, n, n, dx)
        mo.update_mesh_static(TriangleMesh(*sphere))
        if obstacle == "friction":
            mo.friction = 0.25
        sim.add_mesh_obstacle(mo)
    sim.load_marker_particle_data(len(pos), pos.tobytes(), vel.tobytes())
    sim.initialize()
    sim.step(dt)  # seeds rand() with frame + substep (0) before the obstacle's sign tests
    m = sim.get_num_marker_particles()
    assert m == len(pos), "the engine removed particles"
    ep = np.frombuffer(sim.get_marker_particle_position_data_range(0, m), np.float32).reshape(-1, 3)
    solid = None
    if obstacle:
        O.libc_srand(0)
        solid, _ = K.solid_phi_meshes((n, n, n), dx, [sphere])
        assert np.array_equal(solid.view(np.uint32), sim.get_solid_sdf().view(np.uint32))
    if obstacle == "friction":
        with pytest.raises(ValueError):
            Rollout((n, n, n), dx, pos, dt=dt, substeps=1, checkpoint_every=1, tol=1e-9, solid_nodes=solid,
                    solid_friction=mo.friction)
        return
    ro = Rollout((n, n, n), dx, pos, dt=dt, substeps=1, checkpoint_every=1, tol=1e-9, solid_nodes=solid,
                 solid_friction=mo.friction if obstacle else 0.0)
    ro.forward(*init)
    rp = ro.positions()
    moved, fwd_iters, _ = ro.stats()
    say(f"rollout vs engine substep: {len(pos)} particles, {moved} moved by collisions, "
        f"{fwd_iters} PCG iterations, max |diff| = {np.abs(rp - ep).max():.3g}")
    assert np.array_equal(rp, ep)


def _rollout_case(n=16, substeps=3, checkpoint_every=1, dt=1.0 / 240.0):
    dx = 1.0 / n
    pos = _ball(n, 35)
    init = scenes.noise_mac_field(n, seed=36, amplitude=0.2, correlation=4.0)
    target_init = [a + b for a, b in zip(init, scenes.noise_mac_field(n, seed=37, amplitude=0.2, correlation=4.0))]
    ro = Rollout((n, n, n), dx, pos, dt=dt, substeps=substeps, checkpoint_every=checkpoint_every, tol=1e-12,
                 max_iter=2000)
    target = ro.forward(*target_init).astype(np.float64)
    return ro, init, target


def _loss(rho, target):
    return float(((rho.astype(np.float64) - target) ** 2).sum())


def test_rollout_gradient_independent_of_checkpoints():
    grads = []
    for c in (1, 2, 4):
        ro, init, target = _rollout_case(substeps=4, checkpoint_every=c)
        rho = ro.forward(*init)
        g = ro.backward((2.0 * (rho.astype(np.float64) - target)).astype(np.float32))
        again = ro.backward((2.0 * (rho.astype(np.float64) - target)).astype(np.float32))
        assert all(np.array_equal(a, b) for a, b in zip(g, again))
        grads.append(g)
        say(f"checkpoint every {c}: |grad|_max = {max(float(np.abs(a).max()) for a in g):.6g}, stats {ro.stats()}")
        ro.close()
    for g in grads[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(grads[0], g))


def test_rollout_gradient_central_differences():
    ro, init, target = _rollout_case(substeps=3, checkpoint_every=2)
    rho = ro.forward(*init)
    # holding the recorded masks fixed changes nothing at the recorded point
    assert np.array_equal(ro.forward(*init, freeze=True), rho)
    rho = ro.forward(*init)
    g = ro.backward((2.0 * (rho.astype(np.float64) - target)).astype(np.float32))
    moved, fi, bi = ro.stats()
    say(f"rollout: {ro.np} particles, loss {_loss(rho, target):.6g}, {moved} moved by collisions, "
        f"PCG iterations forward {fi}, adjoint {bi}")
    assert moved == 0
    rng = np.random.default_rng(38)
    eps = 1e-2
    # per face: the faces of largest gradient and random faces with nonzero gradient
    flat = [np.abs(a).ravel() for a in g]
    scale = max(float(f.max()) for f in flat)
    picks = []
    for d in range(3):
        order = np.argsort(flat[d])[::-1]
        picks += [(d, int(i)) for i in order[:4]]
        nz = np.flatnonzero(flat[d] > 0.05 * scale)
        picks += [(d, int(i)) for i in rng.choice(nz, size=min(4, len(nz)), replace=False)]
    worst = 0.0
    for d, i in picks:
        plus = [a.copy() for a in init]
        minus = [a.copy() for a in init]
        plus[d].ravel()[i] += eps
        minus[d].ravel()[i] -= eps
        h = float(plus[d].ravel()[i]) - float(minus[d].ravel()[i])
        fd = (_loss(ro.forward(*plus, freeze=True), target) - _loss(ro.forward(*minus, freeze=True), target)) / h
        worst = max(worst, abs(fd - float(g[d].ravel()[i])) / scale)
    say(f"rollout per-face central differences ({len(picks)} faces, eps {eps}): "
        f"max |fd - grad| / max |grad| = {worst:.3g}")
    assert worst <= 2e-2
    # along a smooth direction
    direction = scenes.noise_mac_field(16, seed=39, amplitude=1.0, correlation=4.0)
    t = 1e-2
    lp = _loss(ro.forward(*[a + t * b for a, b in zip(init, direction)], freeze=True), target)
    lm = _loss(ro.forward(*[a - t * b for a, b in zip(init, direction)], freeze=True), target)
    fd = (lp - lm) / (2 * t)
    ad = sum(float(np.dot(a.astype(np.float64).ravel(), b.astype(np.float64).ravel())) for a, b in zip(g, direction))
    say(f"rollout directional derivative: fd {fd:.6g}, adjoint {ad:.6g}")
    assert abs(fd - ad) <= 1e-2 * abs(ad)
    ro.close()


def test_rollout_gradient_linear_in_the_cotangent():
    """The adjoint is linear in the density cotangent at every scale.  A
    cotangent scaled by 2^-40 (an exact scaling; the adjoint's pressure
    right-hand side then lies far below the solver tolerance taken as an
    absolute number) gives the gradient scaled by 2^-40, bit for bit: the
    adjoint's pressure solve is skipped only for a zero right-hand side."""
    ro, init, target = _rollout_case(substeps=2, checkpoint_every=1)
    rho = ro.forward(*init)
    rb = (2.0 * (rho.astype(np.float64) - target)).astype(np.float32)
    g = ro.backward(rb)
    scale = np.float32(2.0 ** -40)
    gs = ro.backward(rb * scale)
    _, _, adjoint_iters = ro.stats()
    say(f"rollout cotangent x 2^-40: |grad|_max {max(float(np.abs(a).max()) for a in gs):.6g}, "
        f"adjoint PCG iterations {adjoint_iters}")
    assert adjoint_iters > 0
    for a, b in zip(g, gs):
        assert np.array_equal((a * scale).view(np.uint32), b.view(np.uint32))
    zero = ro.backward(np.zeros_like(rb))
    assert all(not np.any(a) for a in zero)
    ro.close()


def test_rollout_autograd():
    torch = pytest.importorskip("torch")
    from ffengine import autograd as AG
    ro, init, target = _rollout_case(substeps=2, checkpoint_every=1)
    u, v, w = (torch.tensor(a, requires_grad=True) for a in init)
    rho = AG.RolloutDensity.apply(u, v, w, ro)
    loss = ((rho.double() - torch.from_numpy(target)) ** 2).sum()
    loss.backward()
    want = ro.backward((2.0 * (rho.detach().double().numpy() - target)).astype(np.float32))
    say("rollout autograd: gradient against the entry point")
    for t, a in zip((u, v, w), want):
        assert np.array_equal(t.grad.numpy(), a)
    ro.close()


@pytest.mark.timeout(600)
def test_differentiation_configuration_matches_fp64_cpu_adjoint():
    """The differentiation configuration (dam break at rest, 8 particles per
    cell outside the walls, loss = squared L2 distance of the density after
    32 substeps of 1/240 s to the density of a run from a seeded 0.2 m/s
    perturbation, checkpoints every 8 substeps) at 24^3, against the oracle's
    rollout and its FP64 adjoint:
    - with the reference's dot-product order both forwards take the same
      bits (target and current density);
    - the GPU gradient (float cotangents, FP64 MIC(0)-PCG to 1e-9) equals the
      FP64 CPU adjoint (CG to 1e-13) face by face within 1e-4 of the largest
      gradient, and within 1e-5 in relative L2 norm.  Both take the P2G
      derivative at the forward's float transfer weights and sums: on
      near-empty faces (weight sums of a few 1e-5) those carry relative errors
      of several percent from cancellation in the float kernel, and a double
      re-evaluation of the weights moved the gradient by 4e-4 of its maximum
      after 29 substeps."""
    from scene_util import differentiation_scene
    n, substeps, every, dt, tol = 24, 32, 8, 1.0 / 240.0, 1e-9
    dx = 1.0 / n
    pos, init, target_init = differentiation_scene(n)
    say(f"differentiation configuration {n}^3: {len(pos)} particles, {substeps} substeps")
    ro = Rollout((n, n, n), dx, pos, dt=dt, substeps=substeps, checkpoint_every=every, tol=tol, max_iter=5000)
    ro.set_reference_summation(1)
    orc = O.OracleRollout((n, n, n), dx, pos, dt, substeps, tol=tol, max_iter=5000)
    target = ro.forward(*target_init)
    assert np.array_equal(target.view(np.uint32), orc.forward(*target_init).view(np.uint32))
    rho = ro.forward(*init)
    assert np.array_equal(rho.view(np.uint32), orc.forward(*init).view(np.uint32))
    rb = (2.0 * (rho.astype(np.float64) - target.astype(np.float64))).astype(np.float32)
    g = ro.backward(rb)
    moved, fwd_iters, bwd_iters = ro.stats()
    gc = orc.backward(rb.astype(np.float64))
    scale = max(float(np.abs(a).max()) for a in gc)
    worst = max(float(np.abs(a.astype(np.float64) - b).max()) for a, b in zip(g, gc)) / scale
    num = np.sqrt(sum(float(((a.astype(np.float64) - b) ** 2).sum()) for a, b in zip(g, gc)))
    den = np.sqrt(sum(float((b ** 2).sum()) for b in gc))
    say(f"  loss {float((rb.astype(np.float64) ** 2).sum()) / 4:.6g}, collision moves {moved}, PCG iterations "
        f"forward {fwd_iters} adjoint {bwd_iters}; max |grad_gpu - grad_fp64| / max |grad| = {worst:.3g}, "
        f"relative L2 {num / den:.3g}")
    assert scale > 0
    assert worst <= 1e-4
    assert num / den <= 1e-5
    ro.close()
