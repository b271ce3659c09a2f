"""The oracle's solid update with moving mesh obstacles (oracle/fforacle.cpp
solidUpdate: fluidsimulation.cpp:5219-5426 and the velocity data of
meshlevelset.cpp) against the reference engine built from its sources
(oracle/_ref), which runs its own _updateSolidLevelSet at the same frame
timing through the stage probe.  Compared bit for bit: the nodal solid phi
and the normalized, extrapolated solid face velocities.

Cases: an animated sphere translating across frames beside a static sphere;
an animated sphere with a velocity scale and an inversed animated container;
two islands in one animated mesh, one of them crossing the domain's border;
frame progress clamped at 0 and inside the frame."""
import numpy as np
import pytest

import oracle_lib as O
from ffengine import scenes
from scene_util import SOLID_UPDATE_TIMINGS, frame_progress, moving_obstacle_cases

pytestmark = pytest.mark.skipif(not O.reference_available(), reason="reference build (oracle/_ref) not present")


CASES = moving_obstacle_cases()


def _reference(n, objects, seed, timings):
    n3, dx = (n, n, n), 1.0 / n
    ref = O.ReferenceSim(n3, dx, threads=1)
    ref.load(scenes.dam_break_particles(n, seed=0))
    for o in objects:
        if o.get("animated", True):
            ref.add_animated_obstacle(n3, dx, o["prev"], o["cur"], o["next"], o["tris"], o.get("inversed", False),
                                      o.get("velocity_scale", 1.0))
        else:
            ref.add_mesh_obstacle(n3, dx, o["cur"], o["tris"], o.get("inversed", False))
    ref.initialize()
    out = []
    for frame_dt, remaining, step in timings:
        O.libc_srand(seed)
        ref.update_solid(frame_dt, remaining, step)
        out.append((ref.solid_phi(n3), ref.solid_velocity(n3)))
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_solid_update_matches_reference(case):
    n, seed = 24, 5
    objects = CASES[case]
    timings = SOLID_UPDATE_TIMINGS
    want = _reference(n, objects, seed, timings)
    moving = 0.0
    for (frame_dt, remaining, step), (rphi, rvel) in zip(timings, want):
        progress = frame_progress(frame_dt, remaining, step)
        phi, vel, used = O.solid_update((n, n, n), 1.0 / n, objects, frame_dt, progress, seed)
        assert used > 0
        assert np.array_equal(phi.view(np.uint32), rphi.view(np.uint32)), (case, progress)
        for a, b in zip(vel, rvel):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (case, progress)
        moving = max(moving, max(float(np.abs(a).max()) for a in vel))
    assert moving > 0.1  # the obstacles do move


def _frames(n, frames, make_objects, inversed=None):
    """The dam break at n^3 with mesh obstacles, frames of 1/30 s through the
    reference engine and the oracle; each frame's obstacle meshes come from
    make_objects(frame) as (prev, cur, next, tris, animated); inversed: the
    objects' flags (default none)."""
    dx = 1.0 / n
    n3 = (n, n, n)
    pos = scenes.dam_break_particles(n, seed=1)
    ref = O.ReferenceSim(n3, dx, threads=1)
    ref.load(pos)
    orc = O.OracleSim(n3, dx)
    orc.load(pos)
    handles = []
    objs = make_objects(0)
    inversed = inversed or [False] * len(objs)
    for (prev, cur, nxt, tris, animated), inv in zip(objs, inversed):
        if animated:
            h = ref.add_animated_obstacle(n3, dx, prev, cur, nxt, tris, inv)
        else:
            h = ref.add_mesh_obstacle(n3, dx, cur, tris, inv)
        handles.append((h, orc.add_moving_obstacle(prev, cur, nxt, tris, animated=animated, inversed=inv)))
    ref.initialize()
    orc.initialize()
    # both engines seed rand() with frame + substep at every substep (the
    # reference's _stepFluid does, fluidsimulation.cpp:10079); the oracle
    # through its restatement of the C library's generator
    out = []
    for f in range(frames):
        if f > 0:
            for (h, q), (prev, cur, nxt, tris, animated) in zip(handles, make_objects(f)):
                if animated:
                    ref.update_animated_obstacle(h, prev, cur, nxt, tris)
                    orc.update_moving_obstacle(q, prev, cur, nxt)
        ref.update(1.0 / 30.0)
        orc.update(1.0 / 30.0)
        out.append((ref.particles(), orc.particles(), orc.solid_velocity()))
    return out


def test_frames_with_an_animated_obstacle_match_reference():
    """An animated sphere pushed into the liquid column beside a static one:
    the solid rebuilt at every substep with its velocities, the isolated-pocket
    conditioning, the divergence's solid terms and the velocity constraint.
    Positions and velocities bit-identical to the reference engine for 5
    frames, and the solid velocities are those of the moving sphere."""
    n = 24
    v, t = scenes.sphere_mesh((0.58, 0.25, 0.5), 0.14, 2)
    s, ts = scenes.sphere_mesh((0.2, 0.75, 0.3), 0.1, 1)
    d = np.array([-0.025, 0.005, 0.0], np.float32)

    def objects(f):
        c = v + f * d
        return [(c - d, c, c + d, t, True), (s, s, s, ts, False)]

    frames = _frames(n, 5, objects)
    for f, ((rp, rv), (op, ov), vel) in enumerate(frames):
        assert len(rp) == len(op), f
        assert np.array_equal(rp.view(np.uint32), op.view(np.uint32)), f
        assert np.array_equal(rv.view(np.uint32), ov.view(np.uint32)), f
    speed = 0.025 * 30.0
    assert abs(float(np.abs(frames[-1][2][0]).max()) - speed) < 1e-3 * speed


def test_frames_with_an_inversed_animated_container_match_reference():
    """An animated sphere inside an inversed animated container (the liquid
    held by the container's inside): the inversed group's union, negation and
    velocity grids on every rebuild.  Bit-identical to the reference engine
    for 5 frames."""
    n = 24
    v, t = scenes.sphere_mesh((0.5, 0.3, 0.5), 0.1, 2)
    c, tc = scenes.sphere_mesh((0.5, 0.5, 0.5), 0.62, 2)
    d = np.array([0.0, 0.02, 0.01], np.float32)
    e = np.array([0.008, 0.0, 0.0], np.float32)

    def objects(f):
        a, b = v + f * d, c + f * e
        return [(a - d, a, a + d, t, True), (b - e, b, b + e, tc, True)]

    frames = _frames(n, 5, objects, inversed=[False, True])
    for f, ((rp, rv), (op, ov), vel) in enumerate(frames):
        assert len(rp) == len(op), f
        assert np.array_equal(rp.view(np.uint32), op.view(np.uint32)), f
        assert np.array_equal(rv.view(np.uint32), ov.view(np.uint32)), f
    assert max(float(np.abs(a).max()) for a in frames[-1][2]) > 0.2
