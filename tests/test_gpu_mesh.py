"""Static mesh obstacles on the GPU against the CPU oracle and the reference.

- the fast mesh level set (exact band, crossing-count signs) and the static
  solid SDF with normal, inversed, several, border-crossing and 30-island
  obstacles, bit for bit against the oracle (pinned to the reference build by
  tests/test_oracle_mesh.py), with rand() seeded identically before each;
- the engine with a MeshObject obstacle: its solid SDF after the first
  update equals the oracle's, and whole frames equal the reference engine's
  bit for bit when both take the same rand() seed (the reference's per-frame
  checksums in tests/golden/obstacles_n32.json);
- enable/disable and removal rebuild the solid; animated meshes whose
  topology changes, duplicate adds and unknown removals raise;
- obstacle friction (the constraint's closest-object friction on partly open
  faces) and boundary friction: frames equal the reference engine's bit for
  bit (same fixture).
"""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from ffengine import FluidSimulation, MeshObject, TriangleMesh
from ffengine import kernels as K
from ffengine import scenes

from golden.make_golden import obstacle_cases  # the fixture's obstacle scenes

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("n,center,radius,sub,cells,off", [
    (24, (0.5, 0.45, 0.55), 0.2, 3, None, (0, 0, 0)),
    (40, (0.13, 0.5, 0.5), 0.3, 3, None, (0, 0, 0)),      # crosses the domain border
    (32, (0.6, 0.2, 0.5), 0.15, 4, (14, 11, 17), (5, 7, 3)),  # offset island grid, leaves it
])
def test_mesh_levelset_matches_oracle(n, center, radius, sub, cells, off):
    dx = 1.0 / n
    cells = cells or (n, n, n)
    v, t = scenes.sphere_mesh(center, radius, sub)
    seed = 17 + n
    say(f"mesh level set n={n} cells={cells} off={off} triangles={len(t)}")
    O.libc_srand(seed)
    g, gd = K.mesh_levelset_fast(cells, dx, off, v, t)
    o, od = O.mesh_levelset_fast(cells, dx, off, v, t, 3, seed)
    assert gd == od
    assert np.array_equal(bits(g), bits(o))


def many_spheres(count, seed):
    rng = np.random.default_rng(seed)
    verts, tris, base = [], [], 0
    for _ in range(count):
        v, t = scenes.sphere_mesh(rng.random(3) * 0.7 + 0.15, 0.03 + 0.03 * rng.random(), 1)
        verts.append(v)
        tris.append(t + base)
        base += len(v)
    return np.concatenate(verts), np.concatenate(tris)


@pytest.mark.parametrize("case", ["one", "two", "inversed", "mixed", "islands", "border", "islands_expanded"])
def test_static_solid_matches_oracle(case):
    n = 36
    dx = 1.0 / n
    a = scenes.sphere_mesh((0.55, 0.3, 0.5), 0.17, 3)
    b = scenes.sphere_mesh((0.3, 0.7, 0.35), 0.12, 2)
    c = scenes.sphere_mesh((0.5, 0.5, 0.5), 0.4, 2)
    d = scenes.sphere_mesh((0.95, 0.5, 0.5), 0.2, 2)
    sets = {"one": [a], "two": [a, b], "inversed": [(c[0], c[1], True)], "mixed": [b, (c[0], c[1], True), a],
            "islands": [many_spheres(30, 12)], "border": [d, a],
            "islands_expanded": [many_spheres(30, 12) + (True, 0.02)]}
    seed = 40 + len(case)
    say(f"static solid {case}")
    O.libc_srand(seed)
    g, gd = K.solid_phi_meshes((n, n, n), dx, sets[case])
    o, od = O.solid_phi_meshes((n, n, n), dx, sets[case], seed)
    assert gd == od
    assert np.array_equal(bits(g), bits(o))


def _sim(n, pos, threads=None):
    sim = FluidSimulation(n, n, n, 1.0 / n)
    sim.add_body_force(0.0, -9.81, 0.0)
    sim.PICFLIP_ratio = 0.05
    sim.CFL_condition_number = 5
    sim.enable_surface_reconstruction = False
    if threads:
        sim.max_thread_count = threads
        sim.set_pressure_solver_reference_summation(True)
    sim.load_marker_particle_data(len(pos), pos.tobytes(), np.zeros_like(pos).tobytes())
    return sim


def _obstacle(n, v, t, inversed=False, friction=0.0):
    mo = MeshObject(n, n, n, 1.0 / n)
    mo.update_mesh_static(TriangleMesh(v, t))
    mo.inverse = inversed
    mo.friction = friction
    return mo


def _state(sim):
    m = sim.get_num_marker_particles()
    p = np.frombuffer(sim.get_marker_particle_position_data_range(0, m), np.float32)
    v = np.frombuffer(sim.get_marker_particle_velocity_data_range(0, m), np.float32)
    return p, v


def test_engine_solid_with_obstacle_matches_oracle_and_rebuilds():
    n = 32
    dx = 1.0 / n
    pos = scenes.dam_break_particles(n, seed=3)
    v, t = scenes.sphere_mesh((0.6, 0.25, 0.5), 0.15, 3)
    sim = _sim(n, pos)
    mo = _obstacle(n, v, t)
    sim.add_mesh_obstacle(mo)
    with pytest.raises(RuntimeError):
        sim.add_mesh_obstacle(mo)
    sim.initialize()
    # the engine seeds rand() with frame + substep at every substep, as the
    # reference's _stepFluid does: the first update's rebuild draws after srand(0)
    O.libc_srand(4321)
    sim.update(1.0 / 30.0)
    st = sim.get_frame_stats_data()
    if st.substeps == 1:  # the frame's one substep drew the 6 values of its rebuild after srand(0)
        import ctypes
        assert ctypes.CDLL(None).rand() == int(O.glibc_rand_values(0, 7)[6])
    expect, draws = O.solid_phi_meshes((n, n, n), dx, [(v, t)], 0)
    say(f"engine solid with obstacle: draws {sim.get_last_solid_rand_draws()} (oracle {draws})")
    assert sim.get_last_solid_rand_draws() == draws == 6
    assert np.array_equal(bits(sim.get_solid_sdf()), bits(expect))
    # disabling marks the object changed: the next update rebuilds without it
    mo.enable = False
    sim.update(1.0 / 30.0)
    assert sim.get_last_solid_rand_draws() == 3
    assert np.array_equal(bits(sim.get_solid_sdf()), bits(O.solid_box_phi((n, n, n), dx)))
    mo.enable = True
    sim.update(1.0 / 30.0)  # frame 2: srand(2)
    again, _ = O.solid_phi_meshes((n, n, n), dx, [(v, t)], 2)
    assert np.array_equal(bits(sim.get_solid_sdf()), bits(again))
    sim.remove_mesh_obstacle(mo)
    with pytest.raises(RuntimeError):
        sim.remove_mesh_obstacle(mo)
    sim.update(1.0 / 30.0)
    assert np.array_equal(bits(sim.get_solid_sdf()), bits(O.solid_box_phi((n, n, n), dx)))


def test_unsupported_obstacles_raise():
    n = 16
    v, t = scenes.sphere_mesh((0.5, 0.5, 0.5), 0.2, 1)
    mo = _obstacle(n, v, t)
    with pytest.raises(RuntimeError):
        mo.update_mesh_animated(TriangleMesh(v, t), TriangleMesh(v[:-3], t[:-3]), TriangleMesh(v, t))
    with pytest.raises(ValueError):
        mo.friction = 1.5
    mo.friction = 0.25
    assert abs(mo.friction - 0.25) < 1e-7
    bad = MeshObject(n, n, n, 1.0 / n)
    with pytest.raises(RuntimeError):
        bad.update_mesh_static(TriangleMesh(v, np.array([[0, 1, len(v)]], np.int32)))


def _fixture():
    with open(os.path.join(GOLDEN, "obstacles_n32.json")) as fh:
        return json.load(fh)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _check_frame(sim, want, what):
    pg, vg = _state(sim)
    got = {"count": len(pg) // 3, "positions_sha256": _sha(pg), "velocities_sha256": _sha(vg)}
    say(f"{what}: {got['count']} particles (reference {want['count']}), identical "
        f"{got['positions_sha256'] == want['positions_sha256']}")
    assert got["count"] == want["count"], what
    assert got["positions_sha256"] == want["positions_sha256"], what
    assert got["velocities_sha256"] == want["velocities_sha256"], what


def test_boundary_and_live_friction_match_reference():
    """Boundary friction (the domain object's) and an obstacle's friction
    changed between frames, which the reference reads live: every frame's
    positions and velocities against the reference engine's checksums
    (tests/golden/obstacles_n32.json, written by make_golden.py)."""
    rec = _fixture()["boundary_live_friction"]
    n = 32
    pos = scenes.dam_break_particles(n, seed=rec["seed"])
    assert _sha(pos) == rec["initial_positions_sha256"]
    v, t = scenes.sphere_mesh((0.3, 0.3, 0.5), 0.12, 3)
    sim = _sim(n, pos, threads=_fixture()["threads"])
    mo = _obstacle(n, v, t)
    sim.add_mesh_obstacle(mo)
    sim.boundary_friction = rec["boundary_friction"]
    assert abs(sim.boundary_friction - 0.4) < 1e-12
    with pytest.raises(RuntimeError):
        sim.boundary_friction = 1.5
    sim.initialize()
    for frame, want in enumerate(rec["frames"]):
        if frame == 1:  # the obstacle's friction changes without any other state change
            mo.friction = rec["obstacle_friction_from_frame_1"]
        sim.update(1.0 / 30.0)
        _check_frame(sim, want, f"boundary/live friction frame {frame}")


@pytest.mark.parametrize("case", ["sphere", "inversed_and_sphere", "friction", "inversed_and_sphere_friction"])
def test_engine_frames_with_obstacles_match_reference(case):
    """Whole frames with static mesh obstacles (with and without friction)
    against the reference engine's per-frame checksums."""
    fx = _fixture()
    rec = fx["cases"][case]
    n = 32
    pos = scenes.dam_break_particles(n, seed=rec["seed"])
    assert _sha(pos) == rec["initial_positions_sha256"]
    sim = _sim(n, pos, threads=fx["threads"])
    keep = [_obstacle(n, v, t, inv, fr) for (v, t), inv, fr in obstacle_cases()[case]]
    for mo in keep:
        sim.add_mesh_obstacle(mo)
    sim.initialize()
    for frame, want in enumerate(rec["frames"]):
        sim.update(1.0 / 30.0)
        _check_frame(sim, want, f"obstacles {case} frame {frame}")
