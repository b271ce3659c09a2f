"""The FP64 CPU adjoint of the differentiable rollout (oracle_rollout_* in
oracle/fforacle.h), pinned without the GPU.

The forward is the oracle engine's float substep without particle removal;
the adjoint computes every stage's vector-Jacobian product in double at the
forward's states, with the recorded liquid SDF and P2G flags held fixed and
the collision response taken as the identity.  The pins are the
mathematics' own: central differences of the same map (forwards reusing the
recorded masks) along a smooth direction and per face, and linearity in the
cotangent.  The scene is a ball of liquid that reaches no wall, so the
collision response never acts, with substeps long enough and velocities
large enough that the particles' displacement feeds back into the grids
(with a planted mistake in the oracle, dropping P2G's position cotangent
moves the directional derivative by 5%, dropping the saved field's FLIP
term by 70% and the pressure term by 9%; the bounds below are 0.1% and 1%).
"""
import numpy as np
import pytest

import oracle_lib as O
from ffengine import scenes
from scene_util import ball_particles

N, K, DT, AMP = 12, 4, 1.0 / 60.0, 1.0


@pytest.fixture(scope="module")
def case():
    dx = 1.0 / N
    pos = ball_particles(N, 35)
    init = scenes.noise_mac_field(N, seed=36, amplitude=AMP, correlation=4.0)
    target_init = [a + b for a, b in zip(init, scenes.noise_mac_field(N, seed=37, amplitude=0.2, correlation=4.0))]
    ro = O.OracleRollout((N, N, N), dx, pos, DT, K, tol=1e-12, max_iter=2000)
    target = ro.forward(*target_init).astype(np.float64)
    rho = ro.forward(*init)
    rb = 2.0 * (rho.astype(np.float64) - target)
    g = ro.backward(rb)
    return ro, init, target, rho, rb, g


def _loss(rho, target):
    return float(((rho.astype(np.float64) - target) ** 2).sum())


def test_frozen_masks_reproduce_the_recording_forward(case):
    ro, init, _, rho, _, _ = case
    assert np.array_equal(ro.forward(*init, freeze=True), rho)


def test_directional_derivative_matches_central_differences(case):
    ro, init, target, _, _, g = case
    direction = scenes.noise_mac_field(N, seed=39, amplitude=1.0, correlation=4.0)
    ad = sum(float(np.dot(a.ravel(), b.astype(np.float64).ravel())) for a, b in zip(g, direction))
    errs = []
    for t in (1e-3, 3e-3):
        lp = _loss(ro.forward(*[a + t * b for a, b in zip(init, direction)], freeze=True), target)
        lm = _loss(ro.forward(*[a - t * b for a, b in zip(init, direction)], freeze=True), target)
        errs.append(abs((lp - lm) / (2 * t) - ad) / abs(ad))
    assert min(errs) <= 1e-3, errs


def test_per_face_gradient_matches_central_differences(case):
    ro, init, target, _, _, g = case
    rng = np.random.default_rng(40)
    flat = [np.abs(a).ravel() for a in g]
    scale = max(float(f.max()) for f in flat)
    picks = []
    for d in range(3):
        order = np.argsort(flat[d])[::-1]
        picks += [(d, int(i)) for i in order[:3]]
        nz = np.flatnonzero(flat[d] > 0.05 * scale)
        picks += [(d, int(i)) for i in rng.choice(nz, size=min(3, len(nz)), replace=False)]
    eps = 3e-3
    worst = 0.0
    for d, i in picks:
        plus = [a.copy() for a in init]
        minus = [a.copy() for a in init]
        plus[d].ravel()[i] += eps
        minus[d].ravel()[i] -= eps
        h = float(plus[d].ravel()[i]) - float(minus[d].ravel()[i])
        fd = (_loss(ro.forward(*plus, freeze=True), target) - _loss(ro.forward(*minus, freeze=True), target)) / h
        worst = max(worst, abs(fd - float(g[d].ravel()[i])) / scale)
    assert worst <= 1e-2, worst


def test_adjoint_is_linear_in_the_cotangent(case):
    ro, init, _, rho, rb, g = case
    # a backward differentiates the last forward (the tests above ran others)
    assert np.array_equal(ro.forward(*init), rho)
    # an exact power-of-two scaling: every FP64 operation scales exactly, and
    # the CG's relative stopping test takes the same iterations
    s = 2.0 ** -40
    gs = ro.backward(rb * s)
    for a, b in zip(g, gs):
        assert np.array_equal(a * s, b)
    assert all(not np.any(a) for a in ro.backward(np.zeros_like(rb)))
