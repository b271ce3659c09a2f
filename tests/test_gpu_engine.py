"""End-to-end parity of the HIP engine (FluidSimulation over libffengine.so)
with the CPU oracle stepping the same dam break, plus the engine's error
behaviour through the C ABI.

Tolerances: every kernel keeps the reference's operation order.  With the
PCG dot products in the reference's summation order the engine takes the
reference's bits: the 64^3 dam break matches the reference engine's
checksums for all 60 frames.  With the default parallel-tree dot products
the two runs stay identical until a float pressure value rounds differently;
after that, differences grow with the flow, so those tests bound positions
to 1e-5 dx after one frame and count the particles exactly.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from ffengine import FluidSimulation, scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def make_gpu(n, pos, solid=None, precision="fp64", tol=1e-9):
    sim = FluidSimulation(n, n, n, 1.0 / n)
    sim.add_body_force(0.0, -9.81, 0.0)
    sim.PICFLIP_ratio = 0.05
    sim.CFL_condition_number = 5
    sim.enable_surface_reconstruction = False
    sim.pressure_solver_tolerance = tol
    sim.set_pressure_solver_precision(precision)
    if solid is not None:
        sim.set_solid_phi(solid)
    vel = np.zeros_like(pos)
    sim.load_marker_particle_data(len(pos), pos.tobytes(), vel.tobytes())
    sim.initialize()
    return sim


def gpu_particles(sim):
    n = sim.get_num_marker_particles()
    p = np.frombuffer(sim.get_marker_particle_position_data_range(0, n), np.float32).reshape(-1, 3)
    v = np.frombuffer(sim.get_marker_particle_velocity_data_range(0, n), np.float32).reshape(-1, 3)
    return p, v


@pytest.mark.parametrize("n", [24, 32])
def test_frames_match_oracle(n):
    dx = 1.0 / n
    pos = scenes.dam_break_particles(n, seed=5)
    sim = make_gpu(n, pos)
    orc = O.OracleSim((n, n, n), dx)
    orc.load(pos)
    orc.initialize()
    for frame in range(3):
        say(f"engine n={n} frame={frame}")
        sim.update(1.0 / 30.0)
        orc.update(1.0 / 30.0)
        gp, gv = gpu_particles(sim)
        op, ov = orc.particles()
        st = sim.get_frame_stats_data()
        assert st.substeps == orc.substeps()
        assert len(gp) == len(op)
        err = np.abs(gp - op).max() / dx
        say(f"  particles={len(gp)} substeps={st.substeps} iters={st.pressure_solver_iterations} "
            f"oracle_iters={orc.pressure_iterations()} max|dp|/dx={err:.3g}")
        if frame == 0:
            assert err <= 1e-5


def _golden_dam_break(n, seed):
    import importlib.util
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "make_golden.py")
    spec = importlib.util.spec_from_file_location("make_golden", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.dam_break(n, seed)


def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.timeout(900)
def test_c1_sixty_frames_match_reference_bits():
    """The 64^3 dam break for 60 frames of 1/30 s, FP64 PCG with
    its dot products in the reference's summation order: every frame's
    particle count, positions and velocities hash to what the reference
    engine (built from its sources) produced (tests/golden/e2e_c1_n64.json)."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "e2e_c1_n64.json")
    with open(path) as fh:
        rec = json.load(fh)
    n = 64
    pos = _golden_dam_break(n, 0)
    assert _sha(pos) == rec["initial_positions_sha256"]
    sim = make_gpu(n, pos)
    sim.max_thread_count = 8
    sim.set_pressure_solver_reference_summation(True)
    for fr in rec["frames"]:
        sim.update(1.0 / 30.0)
        p, v = gpu_particles(sim)
        st = sim.get_frame_stats_data()
        same = len(p) == fr["count"] and _sha(p) == fr["positions_sha256"] and _sha(v) == fr["velocities_sha256"]
        say("C1 frame %d: particles %d substeps %d iters %d bit-identical %s" % (
            fr["frame"], len(p), st.substeps, st.pressure_solver_iterations, same))
        assert len(p) == fr["count"], fr["frame"]
        assert _sha(p) == fr["positions_sha256"], fr["frame"]
        assert _sha(v) == fr["velocities_sha256"], fr["frame"]


@pytest.mark.timeout(900)
@pytest.mark.parametrize("precision", ["fp64", "fp32-preconditioner"])
def test_c1_sixty_frames_flow_matches_reference(precision):
    """The 64^3 dam break for 60 frames with the solvers the benchmark times:
    FP64 PCG with the parallel-tree dot products, and FP64 PCG with the FP32
    MIC(0) application.  Their sums are ordered differently from the
    reference's, and the flow amplifies such rounding-level differences, so
    particles cannot be compared one by one after the first frames.  The flow
    observables (particle count, centroid, mean squared speed, occupancy of
    4^3 sub-boxes; scene_util.flow_statistics) are compared with the reference
    engine's every frame (tests/golden/e2e_c1_n64_stats.json).  The fixture
    also holds the oracle run from the same particles with every x coordinate
    moved by one ulp: its distance from the reference is how far the flow
    carries a rounding-level difference.  The GPU runs make such differences
    at every substep rather than once, so theirs grow a few frames earlier:
    the tolerance at frame f is 3x the largest one-ulp distance up to frame
    f + 5, plus a floor (10 particles, 0.01 cells, 1e-4 relative speed, 1e-4
    occupancy)."""
    import json
    import os
    from scene_util import flow_statistics
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "e2e_c1_n64_stats.json")) as fh:
        rec = json.load(fh)
    n = 64
    sim = make_gpu(n, _golden_dam_break(n, 0), precision=precision, tol=1e-9)
    frames = list(zip(rec["reference"], rec["oracle_one_ulp"]))
    env = {"count": [], "centroid": [], "speed": [], "occupancy": []}
    running = {key: 0.0 for key in env}
    for ref, ulp in frames:  # the one-ulp spread up to each frame
        running["count"] = max(running["count"], abs(ref["count"] - ulp["count"]))
        running["centroid"] = max(running["centroid"],
                                  max(abs(a - b) for a, b in zip(ref["centroid"], ulp["centroid"])))
        running["speed"] = max(running["speed"],
                               abs(ref["mean_sq_speed"] - ulp["mean_sq_speed"]) / ref["mean_sq_speed"])
        running["occupancy"] = max(running["occupancy"],
                                   max(abs(a - b) for a, b in zip(ref["occupancy"], ulp["occupancy"])))
        for key in env:
            env[key].append(running[key])
    for f, (ref, ulp) in enumerate(frames):
        sim.update(1.0 / 30.0)
        st = sim.get_frame_stats_data()
        assert st.pressure_solver_success, f + 1
        p, v = gpu_particles(sim)
        got = flow_statistics(p, v, n)
        ahead = min(f + 5, len(frames) - 1)
        spread = {key: env[key][ahead] for key in env}
        d_count = abs(got["count"] - ref["count"])
        d_cent = max(abs(a - b) for a, b in zip(got["centroid"], ref["centroid"]))
        d_speed = abs(got["mean_sq_speed"] - ref["mean_sq_speed"]) / ref["mean_sq_speed"]
        d_occ = max(abs(a - b) for a, b in zip(got["occupancy"], ref["occupancy"]))
        say("C1 %s frame %d: count %d (ref %d), centroid %.3g cells (tol %.3g), speed %.3g (tol %.3g), "
            "occupancy %.3g (tol %.3g)" % (precision, f + 1, got["count"], ref["count"], d_cent,
                                           3 * spread["centroid"] + 0.01, d_speed, 3 * spread["speed"] + 1e-4,
                                           d_occ, 3 * spread["occupancy"] + 1e-4))
        assert d_count <= 3 * spread["count"] + 10, f + 1
        assert d_cent <= 3 * spread["centroid"] + 0.01, f + 1
        assert d_speed <= 3 * spread["speed"] + 1e-4, f + 1
        assert d_occ <= 3 * spread["occupancy"] + 1e-4, f + 1


def test_solid_phi_matches_oracle():
    n = 24
    sim = make_gpu(n, scenes.dam_break_particles(n, seed=1))
    say("engine solid phi")
    assert np.array_equal(sim.get_solid_sdf().view(np.uint32),
                          O.solid_box_phi((n, n, n), 1.0 / n).view(np.uint32))


@pytest.mark.parametrize("precision", ["mixed", "mixed-reliable", "fp32-preconditioner"])
def test_mixed_precision_engine_runs(precision):
    n = 24
    pos = scenes.dam_break_particles(n, seed=2)
    sim = make_gpu(n, pos, precision=precision, tol=1e-8)
    ref = make_gpu(n, pos, precision="fp64", tol=1e-8)
    say("engine %s precision" % precision)
    sim.update(1.0 / 30.0)
    ref.update(1.0 / 30.0)
    stats = sim.get_last_pressure_solve_stats()
    assert stats["success"] and stats["unknowns"] > 0
    p, v = gpu_particles(sim)
    assert np.isfinite(p).all() and np.isfinite(v).all()
    # both solves meet the same FP64 residual test, so the first frame's
    # positions agree with the FP64 run far below a cell
    pr, _ = gpu_particles(ref)
    assert len(p) == len(pr)
    assert np.abs(p - pr).max() <= 1e-4 / n


def test_error_paths():
    n = 16
    say("engine error paths")
    sim = FluidSimulation(n, n, n, 1.0 / n)
    with pytest.raises(RuntimeError):
        sim.update(1.0 / 30.0)  # before initialize
    with pytest.raises(RuntimeError, match="PICFLIP ratio"):
        sim.PICFLIP_ratio = 1.5
    with pytest.raises(RuntimeError, match="CFL"):
        sim.CFL_condition_number = 0
    with pytest.raises(RuntimeError, match="PICAPIC ratio"):
        sim.PICAPIC_ratio = -0.1
    pos = scenes.dam_break_particles(n, seed=1)
    sim.load_marker_particle_data(len(pos), pos.tobytes(), np.zeros_like(pos).tobytes())
    sim.initialize()
    with pytest.raises(RuntimeError, match="delta time"):
        sim.update(-1.0)
    with pytest.raises(RuntimeError, match="invalid range"):
        sim.get_marker_particle_position_data_range(0, sim.get_num_marker_particles() + 1)
    # the non-range velocity getter returns positions, like the reference
    assert sim.get_marker_particle_velocity_data() == sim.get_marker_particle_position_data()


def test_warm_start_meets_the_same_tolerance():
    """A PCG warm-started from the previous substep's pressure stops at the
    same threshold (tolerance x |b|_inf): every solve converges, within the
    same iteration cap, and the particles stay within a small fraction of a
    cell of the cold-started run over three frames."""
    n = 32
    pos = scenes.dam_break_particles(n, seed=31)
    cold = make_gpu(n, pos)
    warm = make_gpu(n, pos)
    warm.set_pressure_solver_warm_start(True)
    assert warm.get_pressure_solver_warm_start() and not cold.get_pressure_solver_warm_start()
    ic, iw = [], []
    for frame in range(3):
        cold.update(1.0 / 30.0)
        warm.update(1.0 / 30.0)
        sc, sw = cold.get_frame_stats_data(), warm.get_frame_stats_data()
        assert sc.pressure_solver_success and sw.pressure_solver_success, frame
        ic.append(sc.pressure_solver_iterations)
        iw.append(sw.pressure_solver_iterations)
        pc, _ = gpu_particles(cold)
        pw, _ = gpu_particles(warm)
        assert pc.shape == pw.shape, frame
        d = float(np.abs(pc - pw).max()) * n
        say(f"warm start frame {frame}: iterations cold {ic[-1]} warm {iw[-1]}, max |dx| = {d:.3g} cells")
        assert d < 1e-2, frame
    assert sum(iw[1:]) < sum(ic[1:])


def test_removal_rules_match_oracle():
    """_removeMarkerParticles (fluidsimulation.cpp:7773-7851) with both of its
    own rules firing, against the oracle (itself pinned to the reference build
    on this scene) bit for bit over two frames: a cell crowded with 300
    particles, which the 250-per-cell cap thins in particle order, and 8
    particles at 100 m/s, above the speed limit of
    _getMarkerParticleSpeedLimit (7723-7771).  No gravity, so the crowd stays
    in its cell until the cap acts."""
    from scene_util import removal_scene
    n = 20
    dx = 1.0 / n
    pos, vel, _, _ = removal_scene(n, seed=5)
    sim = FluidSimulation(n, n, n, dx)
    sim.PICFLIP_ratio = 0.05
    sim.CFL_condition_number = 5
    sim.enable_surface_reconstruction = False
    sim.max_thread_count = 1
    sim.set_pressure_solver_reference_summation(True)
    sim.load_marker_particle_data(len(pos), pos.tobytes(), vel.tobytes())
    sim.initialize()
    orc = O.OracleSim((n, n, n), dx, gravity=(0.0, 0.0, 0.0))
    orc.load(pos, vel)
    orc.initialize()
    limit = 6 * 5 * dx * 30.0  # 6 CFL steps per frame at dt = 1/30
    for frame in range(2):
        sim.update(1.0 / 30.0)
        orc.update(1.0 / 30.0)
        gp, gv = gpu_particles(sim)
        op, ov = orc.particles()
        say(f"removal frame {frame}: {len(gp)} particles (oracle {len(op)}, seeded {len(pos)})")
        assert len(gp) == len(op), frame
        assert np.array_equal(gp.view(np.uint32), op.view(np.uint32)), frame
        assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32)), frame
        assert len(gp) <= len(pos) - 58, frame  # the cap took at least 50, the limit the 8
        assert float(np.sqrt((gv.astype(np.float64) ** 2).sum(1)).max()) < limit, frame


@pytest.mark.parametrize("precision", ["fp64", "mixed", "fp32-preconditioner"])
def test_preconditioner_schedules_give_the_same_bits(precision):
    """Schedule 2 (one launch per band), 4 (the line pipeline) and 5 (the
    bands recorded once per solve as a HIP graph and replayed) apply the same
    operations to the same operands, so two frames of a dam break come out
    bit for bit the same under each, in each PCG mode."""
    n = 32
    pos = scenes.dam_break_particles(n, seed=3)
    out = {}
    for schedule in (2, 4, 5):
        sim = make_gpu(n, pos, precision=precision, tol=1e-6)
        sim.set_pressure_solver_preconditioner_schedule(schedule)
        assert sim.get_pressure_solver_preconditioner_schedule() == schedule
        its = []
        for _ in range(2):
            sim.update(1.0 / 30.0)
            its.append(sim.get_last_pressure_solve_stats()["iterations"])
        say(f"{precision} schedule {schedule}: iterations {its}")
        p, v = gpu_particles(sim)
        out[schedule] = (p.copy(), v.copy(), its)
    for schedule in (4, 5):
        assert out[schedule][2] == out[2][2]
        assert np.array_equal(out[schedule][0].view(np.uint32), out[2][0].view(np.uint32)), schedule
        assert np.array_equal(out[schedule][1].view(np.uint32), out[2][1].view(np.uint32)), schedule
