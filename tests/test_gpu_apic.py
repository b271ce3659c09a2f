"""The APIC velocity transfer on the GPU against the CPU oracle, bit for bit.

The oracle's APIC P2G and G2P are pinned to the reference build by
tests/test_oracle_vs_reference.py (stage probes and whole frames).  Here:
- the GPU P2G (record sort by block-frame cell, per-face merge in particle
  order) against the oracle, on grids whose face counts are and are not
  multiples of the 10^3 blocks, and on a dense cluster;
- the GPU G2P (sample plus the field's gradients) against the oracle,
  including points outside the domain;
- whole engine frames with APIC (positions, velocities, affine vectors);
- affine data loaded with the particles, read back before the first step.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from ffengine import kernels as K
from ffengine import scenes
from ffengine.fluidsimulation import FluidSimulation

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("n,seed", [(20, 1), (33, 2), (48, 3)])
def test_p2g_apic_bitexact(n, seed):
    dx = 1.0 / n
    rng = np.random.default_rng(seed)
    pos = scenes.dam_break_particles(n, seed=seed)
    vel = (rng.standard_normal(pos.shape) * 0.5).astype(np.float32)
    aff = (rng.standard_normal((len(pos), 9)) * 2.0).astype(np.float32)
    r = 0.5 * dx * np.sqrt(3.0)
    say(f"APIC P2G n={n}: {len(pos)} particles")
    (gu, gv, gw), (bu, bv, bw) = K.p2g_apic((n, n, n), dx, pos, vel, aff, r)
    (ou, ov, ow), (cu, cv, cw) = O.p2g_apic((n, n, n), dx, pos, vel, aff, r)
    for g, o in zip((gu, gv, gw), (ou, ov, ow)):
        assert np.array_equal(bits(g), bits(o))
    for g, o in zip((bu, bv, bw), (cu, cv, cw)):
        assert np.array_equal(g, o)


def test_p2g_apic_dense_cluster_bitexact():
    n = 16
    dx = 1.0 / n
    rng = np.random.default_rng(7)
    # 60 particles per cell in a 3x3x3 cell cluster straddling block faces
    base = np.array([9.0, 9.0, 9.0]) * dx
    pos = (base + rng.random((60 * 27, 3)) * 3 * dx).astype(np.float32)
    vel = rng.standard_normal(pos.shape).astype(np.float32)
    aff = rng.standard_normal((len(pos), 9)).astype(np.float32)
    r = 0.5 * dx * np.sqrt(3.0)
    (gu, gv, gw), valid_g = K.p2g_apic((n, n, n), dx, pos, vel, aff, r)
    (ou, ov, ow), valid_o = O.p2g_apic((n, n, n), dx, pos, vel, aff, r)
    for g, o in zip((gu, gv, gw), (ou, ov, ow)):
        assert np.array_equal(bits(g), bits(o))
    for g, o in zip(valid_g, valid_o):
        assert np.array_equal(g, o)


@pytest.mark.parametrize("n,seed", [(17, 4), (40, 5)])
def test_g2p_apic_bitexact(n, seed):
    dx = 1.0 / n
    rng = np.random.default_rng(seed)
    uvw = [rng.standard_normal(s).astype(np.float32) for s in [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)]]
    pts = (rng.random((20000, 3)) * 1.04 - 0.02).astype(np.float32)
    gv, ga = K.g2p_apic((n, n, n), dx, uvw, pts)
    ov, oa = O.g2p_apic((n, n, n), dx, uvw, pts)
    assert np.array_equal(bits(gv), bits(ov))
    assert np.array_equal(bits(ga), bits(oa))


def _apic_sim(n, pos, vel):
    sim = FluidSimulation(n, n, n, 1.0 / n)
    sim.add_body_force(0.0, -9.81, 0.0)
    sim.CFL_condition_number = 5
    sim.enable_surface_reconstruction = False
    sim.set_velocity_transfer_method_APIC()
    assert sim.is_velocity_transfer_method_APIC()
    sim.load_marker_particle_data(len(pos), pos.tobytes(), vel.tobytes())
    sim.initialize()
    return sim


def _gpu_state(sim):
    m = sim.get_num_marker_particles()
    p = np.frombuffer(sim.get_marker_particle_position_data_range(0, m), np.float32).reshape(-1, 3)
    v = np.frombuffer(sim.get_marker_particle_velocity_data_range(0, m), np.float32).reshape(-1, 3)
    a = np.concatenate([np.frombuffer(getattr(sim, "get_marker_particle_affine%s_data_range" % c)(0, m),
                                      np.float32).reshape(-1, 3) for c in "xyz"], axis=1)
    return p, v, a


@pytest.mark.parametrize("n", [24, 32])
def test_engine_apic_frames_match_oracle(n):
    dx = 1.0 / n
    pos = scenes.dam_break_particles(n, seed=9)
    vel = (np.random.default_rng(10).standard_normal(pos.shape) * 0.2).astype(np.float32)
    sim = _apic_sim(n, pos, vel)
    orc = O.OracleSim((n, n, n), dx, apic=True)
    orc.load(pos, vel)
    orc.initialize()
    for frame in range(3):
        say(f"APIC engine n={n} frame {frame}")
        sim.update(1.0 / 30.0)
        orc.update(1.0 / 30.0)
        gp, gv, ga = _gpu_state(sim)
        op, ov = orc.particles()
        oa = orc.affine()
        assert gp.shape == op.shape, frame
        assert np.array_equal(bits(gp), bits(op)), frame
        assert np.array_equal(bits(gv), bits(ov)), frame
        assert np.array_equal(bits(ga), bits(oa)), frame


def test_affine_data_loads_with_particles():
    n = 16
    dx = 1.0 / n
    rng = np.random.default_rng(11)
    pos = (rng.random((500, 3)) * 1.2 - 0.1).astype(np.float32)  # some outside the domain: dropped
    vel = rng.standard_normal(pos.shape).astype(np.float32)
    aff = [rng.standard_normal(pos.shape).astype(np.float32) for _ in range(3)]
    sim = FluidSimulation(n, n, n, dx)
    sim.set_velocity_transfer_method_APIC()
    sim.load_marker_particle_data(len(pos), pos.tobytes(), vel.tobytes())
    sim.load_marker_particle_affine_data(len(pos), *[a.tobytes() for a in aff])
    sim.initialize()
    gp, gv, ga = _gpu_state(sim)
    inside = np.all((pos >= 0) & (pos < 1.0), axis=1)
    assert np.array_equal(gp, pos[inside])
    assert np.array_equal(ga, np.concatenate([a[inside] for a in aff], axis=1))
    sim.PICAPIC_ratio = 0.25
    assert sim.PICAPIC_ratio == 0.25
    with pytest.raises(RuntimeError):
        sim.PICAPIC_ratio = 1.5
