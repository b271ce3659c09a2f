"""Savestates in the add-on's on-disk format (src/addon/bake.py:554-777,
2795-3100): a run resumed from a savestate continues bit for bit like the
uninterrupted run.

- GPU engine to GPU engine, FLIP and APIC;
- across engines: a savestate the reference engine (built from its sources,
  oracle/_ref) writes loads into the GPU engine, and the GPU engine's next
  frame equals the reference's bit for bit (FP64 PCG with the reference's
  summation order); skipped where that build is absent.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle_lib as O
from ffengine import scenes
from ffengine.fluidsimulation import FluidSimulation
from ffengine.savestate import load_savestate, read_savestate_info, write_savestate

pytestmark = pytest.mark.gpu


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


class _ReferenceAdapter:
    """The reference engine's C ABI behind the few Python methods the
    savestate writer uses."""

    def __init__(self, ref, n3, dx):
        self.ref, self.n3, self.dx = ref, n3, dx
        L = ref.L
        L.FluidSimulation_get_current_frame.restype = C.c_int

    def _range(self, name, start, end):
        out = (C.c_char * (12 * (end - start)))()
        err = C.c_int()
        getattr(self.ref.L, name)(self.ref.h, start, end, out, C.byref(err))
        assert err.value == 1
        return bytes(out)

    def get_grid_dimensions(self):
        return self.n3

    def get_cell_size(self):
        return self.dx

    def is_velocity_transfer_method_APIC(self):
        return False

    def get_num_marker_particles(self):
        err = C.c_int()
        return self.ref.L.FluidSimulation_get_num_marker_particles(self.ref.h, C.byref(err))

    def get_current_frame(self):
        err = C.c_int()
        return self.ref.L.FluidSimulation_get_current_frame(self.ref.h, C.byref(err))

    def get_marker_particle_position_data_range(self, start, end):
        return self._range("FluidSimulation_get_marker_particle_position_data_range", start, end)

    def get_marker_particle_velocity_data_range(self, start, end):
        return self._range("FluidSimulation_get_marker_particle_velocity_data_range", start, end)


@pytest.mark.skipif(not O.reference_available(), reason="oracle/_ref reference build absent")
def test_reference_savestate_resumes_on_gpu(tmp_path):
    n = 24
    dx = 1.0 / n
    pos = scenes.dam_break_particles(n, seed=23)
    vel = (np.random.default_rng(24).standard_normal(pos.shape) * 0.2).astype(np.float32)
    ref = O.ReferenceSim((n, n, n), dx, threads=4)
    ref.load(pos, vel)
    ref.initialize()
    for _ in range(2):
        ref.update(1.0 / 30.0)
    info = write_savestate(_ReferenceAdapter(ref, (n, n, n), dx), str(tmp_path), frameno=1)
    say(f"reference savestate: frame_id {info['frame_id']}, {info['num_marker_particles']} particles")
    gpu = _sim(n, threads=4)
    load_savestate(gpu, str(tmp_path))
    gpu.initialize()
    for frame in range(2):
        ref.update(1.0 / 30.0)
        gpu.update(1.0 / 30.0)
        pr, vr = ref.particles()
        pg, vg = _state(gpu)
        same = np.array_equal(pr.reshape(-1).view(np.uint32), pg.view(np.uint32))
        say(f"reference -> GPU resume frame {frame}: identical {same}")
        assert same, frame
        assert np.array_equal(vr.reshape(-1).view(np.uint32), vg.view(np.uint32)), frame
    assert read_savestate_info(str(tmp_path))["marker_particle_position_filedata"] == "marker_particle_position.data"
