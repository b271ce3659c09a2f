"""Moving (animated) mesh obstacles on the GPU against the CPU oracle, whose
restatement is pinned bit for bit to the reference engine built from its
sources (tests/test_oracle_moving_obstacles.py) and to closed forms
(tests/test_oracle_solid_terms.py):

- the solid update with velocity data (FFAMD_solid_update_moving: the mesh
  at the frame progress and its vertex velocities, per-node closest
  triangles, per-face weighted nearest velocities, unions, the inversed
  objects' negated union, normalization and 5 extrapolation layers): solid
  phi, face velocities and rand() draws bit for bit, for static + animated,
  scaled + inversed animated, and two islands crossing the border, at three
  frame timings;
- the cells' solid weights and the isolated-pocket conditioning of the solid
  velocities, bit for bit (the conditioning also on a 64^3 grid whose random
  closed faces make thousands of groups for the union-find);
- the engine (MeshObject_update_mesh_animated every frame, FP64 PCG with the
  reference's dot order for one thread) against the oracle engine: an
  animated sphere pushed into the liquid beside a static sphere, and an
  animated sphere inside an inversed animated container; particle positions
  and velocities bit-identical every frame for 5 frames, and the solid's face
  velocities too.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from ffengine import FluidSimulation, MeshObject, TriangleMesh
from ffengine import kernels as K
from ffengine import scenes
from scene_util import SOLID_UPDATE_TIMINGS, frame_progress, moving_obstacle_cases
from test_oracle_solid_terms import pocket_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("case", list(moving_obstacle_cases()))
def test_solid_update_matches_oracle(case):
    n, seed = 24, 5
    objects = moving_obstacle_cases()[case]
    for frame_dt, remaining, step in SOLID_UPDATE_TIMINGS:
        progress = frame_progress(frame_dt, remaining, step)
        say(f"solid update {case} at {n}^3, progress {progress}")
        O.libc_srand(seed)
        phi, vel, draws = K.solid_update_moving((n, n, n), 1.0 / n, objects, frame_dt, progress)
        ophi, ovel, used = O.solid_update((n, n, n), 1.0 / n, objects, frame_dt, progress, seed)
        assert draws == used
        assert np.array_equal(bits(phi), bits(ophi)), (case, progress)
        for a, b in zip(vel, ovel):
            assert np.array_equal(bits(a), bits(b)), (case, progress)
    assert max(float(np.abs(a).max()) for a in vel) > 0.1


@pytest.mark.parametrize("n", [16, 40])
def test_cell_weights_match_oracle(n):
    dx = 1.0 / n
    box = O.solid_box_phi((n, n, n), dx)
    sphere = scenes.sphere_phi_nodes(n, center=(0.45, 0.4, 0.55), radius=0.27)
    for phi in (box, sphere, scenes.union_phi(box, sphere)):
        got = K.cell_weights((n, n, n), phi)
        assert np.array_equal(bits(got), bits(O.cell_weights((n, n, n), phi)))
        assert np.any((got > 0) & (got < 1))


@pytest.mark.parametrize("n,seed", [(12, 1), (12, 2), (64, 3)])
def test_conditioning_matches_oracle(n, seed):
    phi, wts, vel = pocket_scene(n, seed)
    got = K.condition_solid_velocity((n, n, n), phi, wts, vel)
    want = O.condition_solid_velocity((n, n, n), phi, wts, vel)
    zeroed = 0
    for a, b, c in zip(got, want, vel):
        assert np.array_equal(bits(a), bits(b))
        zeroed += int(np.count_nonzero(a != c))
    say(f"conditioning {n}^3 seed {seed}: {zeroed} faces zeroed")
    assert zeroed > 0


def _gpu_sim(n, pos):
    sim = FluidSimulation(n, n, n, 1.0 / n)
    sim.add_body_force(0.0, -9.81, 0.0)
    sim.PICFLIP_ratio = 0.05
    sim.CFL_condition_number = 5
    sim.enable_surface_reconstruction = False
    sim.max_thread_count = 1
    sim.set_pressure_solver_reference_summation(True)
    sim.load_marker_particle_data(len(pos), pos.tobytes(), np.zeros_like(pos).tobytes())
    return sim


def _run_frames(n, frames, make_objects):
    dx = 1.0 / n
    pos = scenes.dam_break_particles(n, seed=1)
    sim = _gpu_sim(n, pos)
    orc = O.OracleSim((n, n, n), dx)
    orc.load(pos)
    objs = []
    for prev, cur, nxt, tris, animated, inversed in make_objects(0):
        mo = MeshObject(n, n, n, dx)
        if animated:
            mo.update_mesh_animated(TriangleMesh(prev, tris), TriangleMesh(cur, tris), TriangleMesh(nxt, tris))
        else:
            mo.update_mesh_static(TriangleMesh(cur, tris))
        mo.inverse = inversed
        sim.add_mesh_obstacle(mo)
        objs.append((mo, orc.add_moving_obstacle(prev, cur, nxt, tris, animated=animated, inversed=inversed)))
    sim.initialize()
    orc.initialize()
    # both engines seed rand() with frame + substep at every substep (the
    # reference's _stepFluid does, fluidsimulation.cpp:10079); the oracle
    # through its restatement of the C library's generator
    for f in range(frames):
        if f > 0:
            for (mo, q), (prev, cur, nxt, tris, animated, _) in zip(objs, make_objects(f)):
                if animated:
                    mo.update_mesh_animated(TriangleMesh(prev, tris), TriangleMesh(cur, tris),
                                            TriangleMesh(nxt, tris))
                    orc.update_moving_obstacle(q, prev, cur, nxt)
        sim.update(1.0 / 30.0)
        orc.update(1.0 / 30.0)
        m = sim.get_num_marker_particles()
        gp = np.frombuffer(sim.get_marker_particle_position_data_range(0, m), np.float32).reshape(-1, 3)
        gv = np.frombuffer(sim.get_marker_particle_velocity_data_range(0, m), np.float32).reshape(-1, 3)
        op, ov = orc.particles()
        st = sim.get_frame_stats_data()
        say(f"  frame {f + 1}: {m} particles ({len(op)} oracle), {st.substeps} substeps, "
            f"rand draws {orc.rand_used()}")
        assert st.substeps == orc.substeps(), f
        assert len(gp) == len(op), f
        assert np.array_equal(bits(gp), bits(op)), f
        assert np.array_equal(bits(gv), bits(ov)), f
        for a, b in zip(sim.get_solid_velocity_field(), orc.solid_velocity()):
            assert np.array_equal(bits(a), bits(b)), f
    return sim.get_solid_velocity_field()


def test_engine_frames_with_an_animated_obstacle_match_oracle():
    n = 24
    v, t = scenes.sphere_mesh((0.58, 0.25, 0.5), 0.14, 2)
    s, ts = scenes.sphere_mesh((0.2, 0.75, 0.3), 0.1, 1)
    d = np.array([-0.025, 0.005, 0.0], np.float32)

    def objects(f):
        c = v + f * d
        return [(c - d, c, c + d, t, True, False), (s, s, s, ts, False, False)]

    say(f"animated sphere beside a static one, {n}^3")
    vel = _run_frames(n, 5, objects)
    speed = 0.025 * 30.0
    assert abs(float(np.abs(vel[0]).max()) - speed) < 1e-3 * speed


def test_engine_frames_with_an_inversed_animated_container_match_oracle():
    n = 24
    v, t = scenes.sphere_mesh((0.5, 0.3, 0.5), 0.1, 2)
    c, tc = scenes.sphere_mesh((0.5, 0.5, 0.5), 0.62, 2)
    d = np.array([0.0, 0.02, 0.01], np.float32)
    e = np.array([0.008, 0.0, 0.0], np.float32)

    def objects(f):
        a, b = v + f * d, c + f * e
        return [(a - d, a, a + d, t, True, False), (b - e, b, b + e, tc, True, True)]

    say(f"animated sphere in an inversed animated container, {n}^3")
    _run_frames(n, 5, objects)
