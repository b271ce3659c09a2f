"""The adjoint of the pressure projection against the forward it transposes.

For a fixed liquid SDF and fixed face weights the projection u' = P(u) is
linear, so its adjoint is the transpose of its Jacobian.  Pins:
- the full Jacobian of the CPU oracle's FP64 forward (one solve per face at
  10^3) transposed, applied to random cotangents;
- central finite differences of the GPU forward at eps = 1e-3 on 64 random
  faces at 24^3;
- the torch.autograd wrapper's gradient equals the adjoint entry point's.
Tolerances: the forward rounds the pressure to float before the velocity
update, so each Jacobian column carries relative rounding of about 1e-7;
the solves run to 1e-12 relative residual.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from ffengine import kernels as K
from ffengine import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scene(n):
    """Liquid: a slab over 30% of the height joined with a ball above it, so
    the free surface has flat and curved parts; static box walls."""
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    ball = np.sqrt((x - 0.55) ** 2 + (y - 0.45) ** 2 + (z - 0.5) ** 2) - 0.22
    phi = np.minimum(y - 0.3, ball).astype(np.float32)
    weights = O.face_weights((n, n, n), O.solid_box_phi((n, n, n), dx))
    return (n, n, n), dx, 1.0 / 120.0, phi, weights


def _shapes(n3):
    ni, nj, nk = n3
    return [(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]


def _oracle_project(n3, dx, dt, uvw, weights, phi):
    p, _, _, ok, m = O.pressure_solve(n3, dx, dt, uvw, weights, phi, tol=1e-12, max_iter=5000)
    (u, v, w), _ = O.apply_pressure(n3, dx, dt, p, phi, weights, uvw)
    return np.concatenate([u.ravel(), v.ravel(), w.ravel()]).astype(np.float64)


def test_projection_vjp_is_oracle_jacobian_transposed():
    n3, dx, dt, phi, weights = _scene(10)
    shapes = _shapes(n3)
    sizes = [int(np.prod(s)) for s in shapes]
    total = sum(sizes)
    say(f"projection Jacobian: {total} oracle solves")
    J = np.zeros((total, total))
    for f in range(total):
        e = np.zeros(total, np.float32)
        e[f] = 1.0
        parts = np.split(e, np.cumsum(sizes)[:-1])
        J[:, f] = _oracle_project(n3, dx, dt, [p.reshape(s) for p, s in zip(parts, shapes)], weights, phi)
    rng = np.random.default_rng(0)
    for trial in range(3):
        wbar = rng.standard_normal(total).astype(np.float32)
        want = J.T @ wbar.astype(np.float64)
        parts = np.split(wbar, np.cumsum(sizes)[:-1])
        (gu, gv, gw), it, ok = K.projection_vjp(n3, dx, dt, weights, phi, [p.reshape(s) for p, s in zip(parts, shapes)])
        assert ok
        got = np.concatenate([gu.ravel(), gv.ravel(), gw.ravel()]).astype(np.float64)
        err = np.abs(got - want).max() / np.abs(want).max()
        say(f"  trial {trial}: adjoint iterations {it}, max |error| / max |J^T w| = {err:.3g}")
        assert err <= 1e-5


def test_projection_vjp_central_differences():
    n3, dx, dt, phi, weights = _scene(24)
    shapes = _shapes(n3)
    rng = np.random.default_rng(1)
    # P is linear, so its Jacobian is the same at every state; a small state
    # keeps the float rounding of each forward's pressure far below the
    # differences taken (|P u| rounding ~ 6e-8 |u| per face)
    uvw = [(rng.standard_normal(s) * 0.01).astype(np.float32) for s in shapes]
    wbar = [rng.standard_normal(s).astype(np.float32) for s in shapes]

    def loss(fields):
        _, out, _, _, _, ok = K.pressure_solve(n3, dx, dt, fields, weights, phi, tol=1e-12, max_iter=5000, mode=0)
        assert ok
        return sum(float(np.dot(o.ravel().astype(np.float64), wb.ravel().astype(np.float64)))
                   for o, wb in zip(out, wbar))

    (gu, gv, gw), _, ok = K.projection_vjp(n3, dx, dt, weights, phi, wbar)
    assert ok
    grads = [gu, gv, gw]
    eps = 1e-3
    worst = 0.0
    scale = max(float(np.abs(g).max()) for g in grads)
    for t in range(64):
        d = int(rng.integers(3))
        idx = tuple(int(rng.integers(s)) for s in shapes[d])
        plus = [a.copy() for a in uvw]
        minus = [a.copy() for a in uvw]
        plus[d][idx] += eps
        minus[d][idx] -= eps
        # the perturbation actually stored in float32
        h = float(plus[d][idx]) - float(minus[d][idx])
        fd = (loss(plus) - loss(minus)) / h
        worst = max(worst, abs(fd - float(grads[d][idx])) / scale)
    say(f"projection central differences: 64 faces, max |fd - adjoint| / max |adjoint| = {worst:.3g}")
    assert worst <= 1e-3


def test_projection_autograd():
    torch = pytest.importorskip("torch")
    from ffengine import autograd as AG
    n3, dx, dt, phi, weights = _scene(16)
    shapes = _shapes(n3)
    rng = np.random.default_rng(2)
    fields = [torch.tensor((rng.standard_normal(s) * 0.5).astype(np.float32), requires_grad=True) for s in shapes]
    wbar = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    out = AG.pressure_projection(*fields, dx, dt, weights, phi, 1e-12)
    loss = sum((o * torch.from_numpy(wb)).sum() for o, wb in zip(out, wbar))
    loss.backward()
    (gu, gv, gw), _, ok = K.projection_vjp(n3, dx, dt, weights, phi, wbar, tol=1e-12)
    say("projection autograd vs adjoint entry point")
    for f, g in zip(fields, (gu, gv, gw)):
        assert np.array_equal(f.grad.numpy(), g)


def _points(n, count, seed, margin=0.02):
    rng = np.random.default_rng(seed)
    return (rng.random((count, 3)) * (1 + 2 * margin) - margin).astype(np.float32)


def test_p2g_vjp_is_oracle_jacobian_transposed():
    n = 8
    n3, dx = (n, n, n), 1.0 / n
    r = 0.5 * dx * np.sqrt(3.0)
    pos = _points(n, 300, 3, margin=0.0)
    P = len(pos)
    shapes = _shapes(n3)
    say(f"p2g Jacobian: {3 * P} oracle transfers")
    cols = []
    for q in range(P):
        for d in range(3):
            vel = np.zeros((P, 3), np.float32)
            vel[q, d] = 1.0
            (u, v, w), _ = O.p2g(n3, dx, pos, vel, r)
            cols.append(np.concatenate([u.ravel(), v.ravel(), w.ravel()]))
    J = np.stack(cols, 1).astype(np.float64)  # faces x (P*3)
    rng = np.random.default_rng(4)
    sizes = [int(np.prod(s)) for s in shapes]
    for trial in range(3):
        ub = rng.standard_normal(J.shape[0]).astype(np.float32)
        want = (J.T @ ub.astype(np.float64)).reshape(P, 3)
        parts = [p.reshape(s) for p, s in zip(np.split(ub, np.cumsum(sizes)[:-1]), shapes)]
        got = K.p2g_vjp(n3, dx, pos, r, parts).astype(np.float64)
        err = np.abs(got - want).max() / np.abs(want).max()
        say(f"  trial {trial}: max |error| / max |J^T u| = {err:.3g}")
        assert err <= 1e-6


def _p2g64(n3, dx, pos, vel, r, sums=None):
    """The P2G forward in float64 by brute force over (particle, face) pairs:
    u_f = sum w v / sum w where sum w > 1e-6, else sum w v, with
    w = 1 - c1 d^6 + c2 d^4 - c3 d^2 for d < r (velocityadvector.cpp:467-536).
    sums, when a list, receives the weight sums."""
    ni, nj, nk = n3
    pos = pos.astype(np.float64)
    vel = vel.astype(np.float64)
    c1, c2, c3 = 4.0 / 9.0 / r ** 6, 17.0 / 9.0 / r ** 4, 22.0 / 9.0 / r ** 2
    out = []
    for d, shape in enumerate(_shapes(n3)):
        k, j, i = np.meshgrid(*(np.arange(s) for s in shape), indexing="ij")
        fpos = np.stack([i, j, k], -1).reshape(-1, 3).astype(np.float64) * dx
        fpos[:, [e for e in range(3) if e != d]] += 0.5 * dx
        d2 = ((fpos[:, None, :] - pos[None, :, :]) ** 2).sum(-1)
        w = np.where(d2 < r * r, 1.0 - c1 * d2 ** 3 + c2 * d2 ** 2 - c3 * d2, 0.0)
        S, W = w @ vel[:, d], w.sum(1)
        if sums is not None:
            sums.append(W.reshape(shape))
        out.append(np.where(W > 1e-6, S / np.where(W > 1e-6, W, 1.0), S).reshape(shape))
    return out


def test_p2g_position_vjp_central_differences():
    """The position cotangent of P2G against central differences of a float64
    forward, itself pinned to the oracle's float forward."""
    n = 8
    n3, dx = (n, n, n), 1.0 / n
    r = 0.5 * dx * np.sqrt(3.0)
    pos = _points(n, 300, 21, margin=0.0)
    rng = np.random.default_rng(22)
    vel = rng.standard_normal(pos.shape).astype(np.float32)
    sums = []
    f64 = _p2g64(n3, dx, pos, vel, r, sums)
    (u, v, w), _ = O.p2g(n3, dx, pos, vel, r)
    # the float forward's weights lose ~3e-7 absolute to cancellation near the
    # cutoff, which a small weight sum turns into a larger error of u
    top = max(float(np.abs(a).max()) for a in f64)
    pin = max(float(np.abs(a - b).max()) for a, b in zip(f64, (u, v, w))) / top
    heavy = max(float(np.abs(np.where(W >= 0.05, a - b, 0)).max()) for a, b, W in zip(f64, (u, v, w), sums)) / top
    say(f"p2g float64 forward against the oracle: max |diff| / max |u| = {pin:.3g} (weight sum >= 0.05: {heavy:.3g})")
    assert pin <= 1e-4 and heavy <= 1e-5
    ub = [rng.standard_normal(s).astype(np.float32) for s in _shapes(n3)]

    def loss(p):
        return sum(float(np.dot(a.ravel(), b.ravel().astype(np.float64))) for a, b in zip(_p2g64(n3, dx, p, vel, r), ub))

    vbar, xbar = K.p2g_vjp(n3, dx, pos, r, ub, vel=vel)
    assert np.array_equal(vbar, K.p2g_vjp(n3, dx, pos, r, ub))
    h = 1e-6 * dx
    base = pos.astype(np.float64)
    worst = 0.0
    scale = float(np.abs(xbar).max())
    for t in range(48):
        q, e = int(rng.integers(len(pos))), int(rng.integers(3))
        a, b = base.copy(), base.copy()
        a[q, e] += h
        b[q, e] -= h
        worst = max(worst, abs((loss(a) - loss(b)) / (2 * h) - float(xbar[q, e])) / scale)
    dirs = rng.standard_normal(pos.shape)
    fd = (loss(base + h * dirs) - loss(base - h * dirs)) / (2 * h)
    ad = float(np.dot(xbar.astype(np.float64).ravel(), dirs.ravel()))
    say(f"p2g position cotangent: max |fd - adjoint| / max |adjoint| = {worst:.3g}; "
        f"direction {fd:.6g} vs {ad:.6g}")
    assert worst <= 1e-4
    assert abs(fd - ad) <= 1e-4 * np.abs(xbar).astype(np.float64).sum()


def test_sample_vjp_is_oracle_jacobian_transposed():
    n = 6
    n3, dx = (n, n, n), 1.0 / n
    pos = _points(n, 200, 5)
    shapes = _shapes(n3)
    sizes = [int(np.prod(s)) for s in shapes]
    total = sum(sizes)
    say(f"sampling Jacobian: {total} oracle samplings")
    J = np.zeros((3 * len(pos), total))
    for f in range(total):
        e = np.zeros(total, np.float32)
        e[f] = 1.0
        parts = [p.reshape(s) for p, s in zip(np.split(e, np.cumsum(sizes)[:-1]), shapes)]
        J[:, f] = O.sample_velocity(n3, dx, parts, pos).astype(np.float64).ravel()
    rng = np.random.default_rng(6)
    for trial in range(3):
        vb = rng.standard_normal((len(pos), 3)).astype(np.float32)
        want = J.T @ vb.astype(np.float64).ravel()
        got = np.concatenate([a.ravel() for a in K.sample_velocity_vjp(n3, dx, pos, vb)]).astype(np.float64)
        err = np.abs(got - want).max() / np.abs(want).max()
        say(f"  trial {trial}: max |error| / max |J^T v| = {err:.3g}")
        assert err <= 1e-6


def test_transfers_central_differences():
    n = 24
    n3, dx = (n, n, n), 1.0 / n
    r = 0.5 * dx * np.sqrt(3.0)
    pos = scenes.dam_break_particles(n, seed=8)
    rng = np.random.default_rng(9)
    shapes = _shapes(n3)
    eps = 1e-3
    # P2G: L(vel) = <ubar, P2G(vel)>, derivative in 64 random particle velocities
    vel = (rng.standard_normal(pos.shape) * 0.01).astype(np.float32)
    ub = [rng.standard_normal(s).astype(np.float32) for s in shapes]

    def lp(v):
        out, _ = K.p2g(n3, dx, pos, v, r)
        return sum(float(np.dot(o.ravel().astype(np.float64), b.ravel().astype(np.float64))) for o, b in zip(out, ub))

    g = K.p2g_vjp(n3, dx, pos, r, ub)
    worst = 0.0
    for t in range(64):
        q, d = int(rng.integers(len(pos))), int(rng.integers(3))
        a, b = vel.copy(), vel.copy()
        a[q, d] += eps
        b[q, d] -= eps
        h = float(a[q, d]) - float(b[q, d])
        worst = max(worst, abs((lp(a) - lp(b)) / h - float(g[q, d])) / float(np.abs(g).max()))
    say(f"p2g central differences: max |fd - adjoint| / max |adjoint| = {worst:.3g}")
    assert worst <= 1e-3
    # sampling: L(u) = <vbar, sample(u)>, derivative in 64 random faces
    uvw = [(rng.standard_normal(s) * 0.01).astype(np.float32) for s in shapes]
    q = _points(n, 5000, 10)
    vb = rng.standard_normal((len(q), 3)).astype(np.float32)

    def ls(fields):
        return float(np.dot(K.sample_velocity(n3, dx, fields, q).astype(np.float64).ravel(),
                            vb.astype(np.float64).ravel()))

    gs = K.sample_velocity_vjp(n3, dx, q, vb)
    scale = max(float(np.abs(a).max()) for a in gs)
    worst = 0.0
    for t in range(64):
        d = int(rng.integers(3))
        idx = tuple(int(rng.integers(s)) for s in shapes[d])
        plus = [a.copy() for a in uvw]
        minus = [a.copy() for a in uvw]
        plus[d][idx] += eps
        minus[d][idx] -= eps
        h = float(plus[d][idx]) - float(minus[d][idx])
        worst = max(worst, abs((ls(plus) - ls(minus)) / h - float(gs[d][idx])) / scale)
    say(f"sampling central differences: max |fd - adjoint| / max |adjoint| = {worst:.3g}")
    assert worst <= 1e-3


def test_transfer_autograd():
    torch = pytest.importorskip("torch")
    from ffengine import autograd as AG
    n = 16
    n3, dx = (n, n, n), 1.0 / n
    r = 0.5 * dx * np.sqrt(3.0)
    pos = scenes.dam_break_particles(n, seed=11)
    rng = np.random.default_rng(12)
    vel = torch.tensor((rng.standard_normal(pos.shape) * 0.3).astype(np.float32), requires_grad=True)
    u, v, w = AG.P2GTransfer.apply(vel, pos, n3, dx, r)
    samples = AG.SampleVelocity.apply(u, v, w, pos, dx)
    target = torch.from_numpy(rng.standard_normal(pos.shape).astype(np.float32))
    loss = ((samples - target) ** 2).sum()
    loss.backward()
    say("transfer autograd: P2G then sampling, gradient against the chained entry points")
    gvel = 2.0 * (samples.detach().numpy() - target.numpy())
    gu, gv, gw = K.sample_velocity_vjp(n3, dx, pos, gvel.astype(np.float32))
    want = K.p2g_vjp(n3, dx, pos, r, [gu, gv, gw])
    assert np.array_equal(vel.grad.numpy(), want)


def _rk3_restated(n3, dx, uvw, p0, dt):
    """_RK3 (fluidsimulation.cpp:7616-7623) in float32 over the oracle's
    trilinear sampling: k1 = V(p0), k2 = V(p0 + 0.5 dt k1),
    k3 = V(p0 + 0.75 dt k2), p1 = p0 + dt/9 (2 k1 + 3 k2 + 4 k3)."""
    f = np.float32
    k1 = O.sample_velocity(n3, dx, uvw, p0)
    k2 = O.sample_velocity(n3, dx, uvw, p0 + f(0.5 * dt) * k1)
    k3 = O.sample_velocity(n3, dx, uvw, p0 + f(0.75 * dt) * k2)
    return p0 + f(dt / float(np.float32(9.0))) * ((f(2.0) * k1 + f(3.0) * k2) + f(4.0) * k3)


def _linear_field(n3, dx, A, b):
    """MAC face grids sampling the affine velocity V(x) = A x + b exactly."""
    ni, nj, nk = n3
    out = []
    for d, shape in enumerate(_shapes(n3)):
        k, j, i = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
        x = [i * dx, j * dx, k * dx]
        for e in range(3):
            if e != d:
                x[e] = x[e] + 0.5 * dx
        out.append((A[d, 0] * x[0] + A[d, 1] * x[1] + A[d, 2] * x[2] + b[d]).astype(np.float32))
    return out


def test_advect_rk3_matches_restatement():
    n = 24
    n3, dx, dt = (n, n, n), 1.0 / n, 1.0 / 240.0
    rng = np.random.default_rng(13)
    uvw = [(rng.standard_normal(s) * 1.5).astype(np.float32) for s in _shapes(n3)]
    pos = _points(n, 20000, 14)
    got = K.advect_rk3(n3, dx, uvw, pos, dt)
    want = _rk3_restated(n3, dx, uvw, pos, dt)
    say("advect RK3 forward against the float32 restatement over the oracle's sampling")
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_advect_rk3_vjp_central_differences():
    n = 24
    n3, dx, dt = (n, n, n), 1.0 / n, 1.0 / 240.0
    rng = np.random.default_rng(15)
    # positions: an affine field has no kinks, so differences across cell
    # faces stay exact; points keep 3 cells from the walls
    A = rng.standard_normal((3, 3)) * 2.0
    b = rng.standard_normal(3) * 0.5
    lin = _linear_field(n3, dx, A, b)
    pos = (3 * dx + rng.random((4000, 3)) * (1 - 6 * dx)).astype(np.float32)
    xb = rng.standard_normal(pos.shape).astype(np.float32)
    pbar, _ = K.advect_rk3_vjp(n3, dx, lin, pos, dt, xb)

    def lpos(p):
        return float(np.dot(K.advect_rk3(n3, dx, lin, p, dt).astype(np.float64).ravel(), xb.astype(np.float64).ravel()))

    eps = 1e-3
    worst = 0.0
    for t in range(64):
        q, e = int(rng.integers(len(pos))), int(rng.integers(3))
        a, c = pos.copy(), pos.copy()
        a[q, e] += eps
        c[q, e] -= eps
        h = float(a[q, e]) - float(c[q, e])
        worst = max(worst, abs((lpos(a) - lpos(c)) / h - float(pbar[q, e])) / float(np.abs(pbar).max()))
    say(f"advect d/dpositions: max |fd - adjoint| / max |adjoint| = {worst:.3g}")
    assert worst <= 1e-3
    # grids: a random field, 64 random faces.  A face moves the new positions
    # by about dt * eps * (trilinear weight); the step and eps are sized so
    # that this stays far above the float32 spacing of the positions (6e-8)
    dtg, eps = 1.0 / 24.0, 1e-2
    uvw = [(rng.standard_normal(s) * 1.5).astype(np.float32) for s in _shapes(n3)]
    pos = _points(n, 4000, 16)
    _, grads = K.advect_rk3_vjp(n3, dx, uvw, pos, dtg, xb)

    def lgrid(fields):
        return float(np.dot(K.advect_rk3(n3, dx, fields, pos, dtg).astype(np.float64).ravel(),
                            xb.astype(np.float64).ravel()))

    scale = max(float(np.abs(gg).max()) for gg in grads)
    worst = 0.0
    shapes = _shapes(n3)
    for t in range(64):
        d = int(rng.integers(3))
        idx = tuple(int(rng.integers(s)) for s in shapes[d])
        plus = [x.copy() for x in uvw]
        minus = [x.copy() for x in uvw]
        plus[d][idx] += eps
        minus[d][idx] -= eps
        h = float(plus[d][idx]) - float(minus[d][idx])
        worst = max(worst, abs((lgrid(plus) - lgrid(minus)) / h - float(grads[d][idx])) / scale)
    say(f"advect d/dgrid: max |fd - adjoint| / max |adjoint| = {worst:.3g}")
    assert worst <= 1e-3


def test_advect_autograd():
    torch = pytest.importorskip("torch")
    from ffengine import autograd as AG
    n = 16
    n3, dx, dt = (n, n, n), 1.0 / n, 1.0 / 240.0
    rng = np.random.default_rng(17)
    fields = [torch.tensor((rng.standard_normal(s) * 1.5).astype(np.float32), requires_grad=True)
              for s in _shapes(n3)]
    pos0 = torch.tensor(_points(n, 3000, 18), requires_grad=True)
    p1 = AG.AdvectRK3.apply(pos0, *fields, dx, dt)
    p2 = AG.AdvectRK3.apply(p1, *fields, dx, dt)
    wb = torch.from_numpy(rng.standard_normal(tuple(p2.shape)).astype(np.float32))
    (p2 * wb).sum().backward()
    say("advect autograd: two steps against the chained entry points")
    uvw = [f.detach().numpy() for f in fields]
    pb2, g2 = K.advect_rk3_vjp(n3, dx, uvw, p1.detach().numpy(), dt, wb.numpy())
    pb1, g1 = K.advect_rk3_vjp(n3, dx, uvw, pos0.detach().numpy(), dt, pb2)
    assert np.array_equal(pos0.grad.numpy(), pb1)
    for f, a, c in zip(fields, g1, g2):
        assert np.allclose(f.grad.numpy(), a + c, rtol=0, atol=1e-6 * max(np.abs(a + c).max(), 1e-30))
