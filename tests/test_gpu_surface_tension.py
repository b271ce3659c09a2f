"""Surface tension on the GPU against the CPU oracle, whose restatement is
pinned bit for bit to the reference built from its sources and to closed
forms (tests/test_oracle_surface_tension.py).

- The curvature grid (FFAMD_curvature_grid): the reinitialized surface phi
  and the extrapolated curvature, bit for bit.
- The cluster status (FFAMD_tension_cluster_status): the BAD clusters (the
  only marks the tension terms read), the border and the one-phase blocks,
  exactly; the reference's searches also mark OK cells that stay 0 here.
- The pressure solve with the tension terms (FFAMD_pressure_solve_tension,
  FP64 PCG with the reference's dot order): pressure, velocities and valid
  flags bit for bit, at two constants.
- Whole frames (FluidSimulation_set_surface_tension at two constants, FP64
  PCG in the reference's dot order): particle positions and velocities
  bit-identical to the oracle engine's for 3 frames.  The smooth kernel
  (twice the particle radius) raises on the GPU.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from ffengine import FluidSimulation
from ffengine import kernels as K
from scene_util import ball_particles
from tension_scenes import random_field, tension_scene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("n,drops", [(24, 0), (24, 40), (40, 40), (64, 80)])
def test_curvature_grid_matches_oracle(n, drops):
    say(f"curvature grid {n}^3, {drops} droplets")
    n3, dx, solid, phi = tension_scene(n, drops=drops)
    surf, kg = K.curvature_grid(n3, dx, phi)
    osurf, okg = O.curvature_grid(n3, dx, phi)
    assert np.array_equal(bits(surf), bits(osurf))
    assert np.array_equal(bits(kg), bits(okg))
    assert np.count_nonzero(kg) > 100


@pytest.mark.parametrize("n,seed,drops", [(24, 3, 40), (24, 4, 60), (48, 5, 200)])
def test_cluster_status_matches_oracle(n, seed, drops):
    n3, dx, solid, phi = tension_scene(n, seed=seed, drops=drops)
    got = K.tension_cluster_status(n3, phi)
    want = O.tension_cluster_status(n3, phi)
    assert np.array_equal(got & 12, want & 12)  # BAD inside / outside
    assert np.array_equal(got == 16, want == 16)
    preset = (got == 1) | (got == 2)
    assert np.array_equal(got[preset], want[preset])
    assert np.all(np.isin(got[~preset & (got != 16) & ((got & 12) == 0)], [0]))
    say(f"clusters {n}^3 seed {seed}: {int(np.count_nonzero(got & 12))} BAD cells")
    assert np.count_nonzero(got & 12) > 0


@pytest.mark.parametrize("k", [0.05, 2.0])
def test_pressure_with_tension_matches_oracle(k):
    n = 24
    n3, dx, solid, phi = tension_scene(n)
    wts = O.face_weights(n3, solid)
    uvw = random_field(n)
    p, vel, flags, it, err, ok = K.pressure_solve_tension(n3, dx, 1.0 / 60.0, uvw, wts, phi, k)
    op, ovel, oflags, oit, oerr, ook = O.pressure_solve_tension(n3, dx, 1.0 / 60.0, uvw, wts, phi, k)
    assert ok and ook and it == oit
    assert np.array_equal(bits(p), bits(op))
    for a, b in zip(vel, ovel):
        assert np.array_equal(bits(a), bits(b))
    for a, b in zip(flags, oflags):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("k,smooth", [(0.5, False), (5.0, False)])
def test_frames_with_surface_tension_match_oracle(k, smooth):
    n = 24
    dx = 1.0 / n
    pos = ball_particles(n, 0)
    sim = FluidSimulation(n, n, n, dx)
    sim.add_body_force(0.0, -9.81, 0.0)
    sim.PICFLIP_ratio = 0.05
    sim.CFL_condition_number = 5
    sim.enable_surface_reconstruction = False
    sim.max_thread_count = 1
    sim.set_pressure_solver_reference_summation(True)
    sim.surface_tension = k
    sim.enable_smooth_surface_tension_kernel = smooth
    sim.load_marker_particle_data(len(pos), pos.tobytes(), np.zeros_like(pos).tobytes())
    sim.initialize()
    orc = O.OracleSim((n, n, n), dx)
    orc.load(pos)
    orc.set_surface_tension(k, smooth=smooth)
    orc.initialize()
    for f in range(3):
        sim.update(1.0 / 30.0)
        orc.update(1.0 / 30.0)
        m = sim.get_num_marker_particles()
        gp = np.frombuffer(sim.get_marker_particle_position_data_range(0, m), np.float32).reshape(-1, 3)
        gv = np.frombuffer(sim.get_marker_particle_velocity_data_range(0, m), np.float32).reshape(-1, 3)
        op, ov = orc.particles()
        st = sim.get_frame_stats_data()
        say(f"  tension {k} smooth {smooth} frame {f + 1}: {m} particles, {st.substeps} substeps")
        assert st.substeps == orc.substeps()
        assert len(gp) == len(op), f
        assert np.array_equal(bits(gp), bits(op)), f
        assert np.array_equal(bits(gv), bits(ov)), f


def test_smooth_kernel_raises():
    n = 16
    pos = ball_particles(n, 0)
    sim = FluidSimulation(n, n, n, 1.0 / n)
    sim.surface_tension = 1.0
    sim.enable_smooth_surface_tension_kernel = True
    sim.load_marker_particle_data(len(pos), pos.tobytes(), np.zeros_like(pos).tobytes())
    sim.initialize()
    with pytest.raises(RuntimeError, match="smooth surface tension kernel"):
        sim.update(1.0 / 30.0)
