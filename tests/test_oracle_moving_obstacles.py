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

pytestmark = pytest.mark.skipif(not O.reference_available(), reason="reference build (oracle/_ref) not present")


def _moving(center, radius, sub, d, **kw):
    v, t = scenes.sphere_mesh(center, radius, sub)
    d = np.asarray(d, np.float32)
    return dict(prev=v - d, cur=v, next=v + d, tris=t, **kw)


def _static(center, radius, sub, **kw):
    v, t = scenes.sphere_mesh(center, radius, sub)
    return dict(prev=v, cur=v, next=v, tris=t, animated=False, **kw)


def _two_islands(d):
    a, ta = scenes.sphere_mesh((0.3, 0.3, 0.3), 0.12, 1)
    b, tb = scenes.sphere_mesh((0.05, 0.7, 0.6), 0.15, 1)  # crosses the x = 0 border
    v = np.concatenate([a, b])
    t = np.concatenate([ta, tb + len(a)])
    d = np.asarray(d, np.float32)
    return dict(prev=v - d, cur=v, next=v + 2 * d, tris=t)


CASES = {
    "animated-and-static": [_moving((0.45, 0.35, 0.5), 0.15, 2, (0.02, 0.01, -0.015)),
                            _static((0.7, 0.6, 0.4), 0.1, 1)],
    "scaled-and-inversed": [_moving((0.5, 0.4, 0.5), 0.12, 2, (0.0, -0.03, 0.01), velocity_scale=1.5),
                            _moving((0.5, 0.5, 0.5), 0.6, 2, (0.01, 0.0, 0.0), inversed=True)],
    "islands-across-border": [_two_islands((0.015, -0.01, 0.02))],
}


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
    timings = [(1 / 30, 1 / 30, 1 / 120), (1 / 30, 0.5 / 30, 1 / 240), (1 / 24, 1 / 96, 1 / 96)]
    want = _reference(n, objects, seed, timings)
    moving = 0.0
    for (frame_dt, remaining, step), (rphi, rvel) in zip(timings, want):
        progress = float(np.float32(1.0) - np.float32(remaining + step) / np.float32(frame_dt))
        phi, vel, used = O.solid_update((n, n, n), 1.0 / n, objects, frame_dt, progress, seed)
        assert used > 0
        assert np.array_equal(phi.view(np.uint32), rphi.view(np.uint32)), (case, progress)
        for a, b in zip(vel, rvel):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (case, progress)
        moving = max(moving, max(float(np.abs(a).max()) for a in vel))
    assert moving > 0.1  # the obstacles do move
