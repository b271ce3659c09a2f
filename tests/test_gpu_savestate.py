"""Savestates in the add-on's on-disk format (src/addon/bake.py:554-777,
2795-3100): a run resumed from a savestate continues bit for bit like the
uninterrupted run.

- GPU engine to GPU engine, FLIP and APIC;
- across engines: a savestate the reference engine wrote (a committed
  fixture) loads into the GPU engine, and the GPU engine's next frames equal
  the reference's bit for bit (FP64 PCG with the reference's summation
  order).
"""
import hashlib
import os
import sys

import numpy as np
import pytest

from ffengine import scenes
from ffengine.fluidsimulation import FluidSimulation
from ffengine.savestate import load_savestate, read_savestate_info, write_savestate

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sim(n, apic=False, threads=None):
    sim = FluidSimulation(n, n, n, 1.0 / n)
    sim.add_body_force(0.0, -9.81, 0.0)
    sim.PICFLIP_ratio = 0.05
    sim.CFL_condition_number = 5
    sim.enable_surface_reconstruction = False
    if apic:
        sim.set_velocity_transfer_method_APIC()
    if threads:
        sim.max_thread_count = threads
        sim.set_pressure_solver_reference_summation(True)
    return sim


def _state(sim):
    m = sim.get_num_marker_particles()
    p = np.frombuffer(sim.get_marker_particle_position_data_range(0, m), np.float32)
    v = np.frombuffer(sim.get_marker_particle_velocity_data_range(0, m), np.float32)
    return p, v


@pytest.mark.parametrize("apic", [False, True])
def test_resume_matches_uninterrupted_run(tmp_path, apic):
    n = 24
    pos = scenes.dam_break_particles(n, seed=21)
    vel = (np.random.default_rng(22).standard_normal(pos.shape) * 0.2).astype(np.float32)
    a = _sim(n, apic)
    a.load_marker_particle_data(len(pos), pos.tobytes(), vel.tobytes())
    a.initialize()
    for _ in range(2):
        a.update(1.0 / 30.0)
    info = write_savestate(a, str(tmp_path), frameno=1)
    assert info["frame_id"] == 1 and info["num_marker_particles"] == a.get_num_marker_particles()
    assert bool(info["marker_particle_affinex_filedata"]) == apic
    b = _sim(n, apic)
    load_savestate(b, str(tmp_path))
    b.initialize()
    assert b.get_current_frame() == 2
    for frame in range(2):
        a.update(1.0 / 30.0)
        b.update(1.0 / 30.0)
        pa, va = _state(a)
        pb, vb = _state(b)
        say(f"resume ({'APIC' if apic else 'FLIP'}) frame {frame}: {len(pa) // 3} particles, "
            f"identical {np.array_equal(pa.view(np.uint32), pb.view(np.uint32))}")
        assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)), frame
        assert np.array_equal(va.view(np.uint32), vb.view(np.uint32)), frame
        if apic:
            for c in "xyz":
                g = "get_marker_particle_affine%s_data_range" % c
                m = a.get_num_marker_particles()
                assert getattr(a, g)(0, m) == getattr(b, g)(0, m), frame


def test_reference_savestate_resumes_on_gpu(tmp_path):
    """A savestate the reference engine wrote (tests/golden/savestate_ref_n16.npz,
    make_golden.py: the 16^3 dam break with random velocities after 2 frames)
    loads into the GPU engine, whose next 2 frames equal the reference's bit
    for bit (their SHA-256 in the fixture)."""
    fx = np.load(os.path.join(GOLDEN, "savestate_ref_n16.npz"))
    n = int(fx["n"])
    for key, name in (("position", "marker_particle_position.data"), ("velocity", "marker_particle_velocity.data"),
                      ("state", "autosave.state")):
        (tmp_path / name).write_bytes(fx[key].tobytes())
    info = read_savestate_info(str(tmp_path))
    say(f"reference savestate: frame_id {info['frame_id']}, {info['num_marker_particles']} particles")
    gpu = _sim(n, threads=int(fx["threads"]))
    load_savestate(gpu, str(tmp_path))
    gpu.initialize()
    assert gpu.get_current_frame() == info["frame_id"] + 1
    for frame, (want_p, want_v) in enumerate(fx["next_frames_sha256"]):
        gpu.update(1.0 / 30.0)
        pg, vg = _state(gpu)
        same = hashlib.sha256(pg.tobytes()).hexdigest() == want_p
        say(f"reference -> GPU resume frame {frame}: identical {same}")
        assert same, frame
        assert hashlib.sha256(vg.tobytes()).hexdigest() == want_v, frame
    assert info["marker_particle_position_filedata"] == "marker_particle_position.data"
