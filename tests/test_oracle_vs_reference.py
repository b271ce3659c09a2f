"""The CPU oracle against the reference engine built from its sources
(oracle/_ref, `make -C oracle ref`), on inputs other than the committed golden
ones.  Skipped where that build is absent (it needs the reference sources).
Every comparison is bit-exact.
"""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.skipif(not O.reference_available(), reason="oracle/_ref reference build absent")

_spec = importlib.util.spec_from_file_location(
    "make_golden", os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "make_golden.py"))
G = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(G)
fp = O.fp


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    b = np.ascontiguousarray(b, np.float32).reshape(-1)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("n,seed", [(24, 21), (31, 22)])
def test_stages_match_reference(n, seed):
    P = O.probe()
    dx = 1.0 / n
    n3 = (n, n, n)
    pos = G.dam_break(n, seed)
    vel = (np.random.default_rng(seed + 100).standard_normal(pos.shape) * 0.4).astype(np.float32)
    r = 0.5 * dx * np.sqrt(3.0)

    box = np.zeros((n + 1) ** 3, np.float32)
    P.refprobe_solid_box_phi(n, n, n, C.c_double(dx), fp(box))
    assert bits_equal(O.solid_box_phi(n3, dx), box)
    solid = np.minimum(box, G.sphere_phi(n, (0.55, 0.25, 0.45), 0.17).reshape(-1))
    solid3 = solid.reshape(n + 1, n + 1, n + 1)

    ref = np.zeros(n ** 3, np.float32)
    P.refprobe_liquid_sdf(n, n, n, C.c_double(dx), len(pos), fp(pos), C.c_double(r), fp(ref))
    raw = O.liquid_sdf(n3, dx, pos, r)
    assert bits_equal(raw, ref)
    P.refprobe_liquid_sdf_postprocess(n, n, n, C.c_double(dx), fp(solid), fp(ref))
    phi = O.liquid_sdf_postprocess(n3, dx, solid3, raw)
    assert bits_equal(phi, ref)

    fs = (n + 1) * n * n
    ru = [np.zeros(fs, np.float32) for _ in range(3)]
    rv = [np.zeros(fs, np.uint8) for _ in range(3)]
    P.refprobe_p2g(n, n, n, C.c_double(dx), len(pos), fp(pos), fp(vel), C.c_double(r),
                   *[fp(a) for a in ru], *[fp(a) for a in rv])
    fields, valids = O.p2g(n3, dx, pos, vel, r)
    dims = [(n + 1, n, n), (n, n + 1, n), (n, n, n + 1)]
    ext = []
    for d in range(3):
        assert bits_equal(fields[d], ru[d]) and np.array_equal(valids[d].reshape(-1), rv[d]), d
        P.refprobe_extrapolate(*dims[d], fp(ru[d]), fp(rv[d]), 12)
        e = O.extrapolate(fields[d], valids[d], 12)
        assert bits_equal(e, ru[d]), d
        ext.append(e)

    rw = [np.zeros(fs, np.float32) for _ in range(3)]
    wc = np.zeros(n ** 3, np.float32)
    P.refprobe_face_weights(n, n, n, C.c_double(dx), fp(solid), *[fp(a) for a in rw], fp(wc))
    wts = O.face_weights(n3, solid3)
    for d in range(3):
        assert bits_equal(wts[d], rw[d]), d

    dt = 0.004
    ext[1] -= np.float32(9.81 * dt)
    ref_uvw = [a.reshape(-1).copy() for a in ext]
    pr = np.zeros(n ** 3, np.float32)
    it, err = C.c_int(), C.c_float()
    vv = [np.zeros(fs, np.uint8) for _ in range(3)]
    P.refprobe_pressure_solve(n, n, n, C.c_double(dx), C.c_double(dt), *[fp(a) for a in ref_uvw],
                              *[fp(a) for a in rw], fp(wc), fp(ref), C.c_double(1e-9), 900, fp(pr), C.byref(it),
                              C.byref(err), *[fp(a) for a in vv])
    p, iters, e, ok, m = O.pressure_solve(n3, dx, dt, ext, wts, phi, tol=1e-9)
    assert iters == it.value and np.float32(e) == np.float32(err.value)
    assert bits_equal(p, pr)
    out, valid = O.apply_pressure(n3, dx, dt, p.reshape(n, n, n), phi, wts, ext)
    for d in range(3):
        assert bits_equal(out[d], ref_uvw[d]) and np.array_equal(valid[d].reshape(-1), vv[d]), d

    q = (np.random.default_rng(seed + 200).random((3000, 3)) * 1.02 - 0.01).astype(np.float32)
    sr = np.zeros_like(q)
    P.refprobe_sample_velocity(n, n, n, C.c_double(dx), *[fp(a) for a in ref_uvw], len(q), fp(q), fp(sr))
    assert bits_equal(O.sample_velocity(n3, dx, out, q), sr)


@pytest.mark.parametrize("n,frames", [(24, 4)])
def test_engine_matches_reference(n, frames):
    dx = 1.0 / n
    pos = G.dam_break(n, 3)
    rng = np.random.default_rng(4)
    vel = (rng.standard_normal(pos.shape) * 0.2).astype(np.float32)
    ref = O.ReferenceSim((n, n, n), dx, threads=4)
    ref.load(pos, vel)
    ref.initialize()
    orc = O.OracleSim((n, n, n), dx)
    orc.load(pos, vel)
    orc.initialize()
    for f in range(frames):
        ref.update(1.0 / 30.0)
        orc.update(1.0 / 30.0)
        pr, vr = ref.particles()
        po, vo = orc.particles()
        assert bits_equal(po, pr) and bits_equal(vo, vr), f


@pytest.mark.parametrize("n,seed", [(24, 41), (31, 42)])
def test_apic_transfers_match_reference(n, seed):
    """APIC P2G (velocityadvector.cpp:543-623) and APIC G2P
    (fluidsimulation.cpp:6709-6769, 6791-6843) against the reference's own
    VelocityAdvector and FluidSimulation, bit for bit."""
    P = O.probe()
    dx = 1.0 / n
    n3 = (n, n, n)
    rng = np.random.default_rng(seed)
    pos = G.dam_break(n, seed)
    vel = (rng.standard_normal(pos.shape) * 0.4).astype(np.float32)
    aff = (rng.standard_normal((len(pos), 9)) * 3.0).astype(np.float32)
    r = 0.5 * dx * np.sqrt(3.0)
    fs = (n + 1) * n * n
    ru = [np.zeros(fs, np.float32) for _ in range(3)]
    rv = [np.zeros(fs, np.uint8) for _ in range(3)]
    P.refprobe_p2g_apic(n, n, n, C.c_double(dx), len(pos), fp(pos), fp(vel), fp(aff), C.c_double(r),
                        *[fp(a) for a in ru], *[fp(a) for a in rv])
    fields, valids = O.p2g_apic(n3, dx, pos, vel, aff, r)
    for d in range(3):
        assert bits_equal(fields[d], ru[d]), d
        assert np.array_equal(valids[d].reshape(-1), rv[d]), d
    uvw = [(rng.standard_normal(s) * 0.5).astype(np.float32) for s in [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)]]
    q = (rng.random((3000, 3)) * 1.02 - 0.01).astype(np.float32)
    rvel = np.zeros((len(q), 3), np.float32)
    raff = np.zeros((len(q), 9), np.float32)
    P.refprobe_g2p_apic(n, n, n, C.c_double(dx), *[fp(a) for a in uvw], len(q), fp(q), fp(rvel), fp(raff))
    ovel, oaff = O.g2p_apic(n3, dx, uvw, q)
    assert bits_equal(ovel, rvel)
    assert bits_equal(oaff, raff)


@pytest.mark.parametrize("n,frames", [(24, 4)])
def test_engine_apic_matches_reference(n, frames):
    """Whole frames with the APIC transfer, oracle against the reference."""
    dx = 1.0 / n
    pos = G.dam_break(n, 5)
    rng = np.random.default_rng(6)
    vel = (rng.standard_normal(pos.shape) * 0.2).astype(np.float32)
    ref = O.ReferenceSim((n, n, n), dx, threads=4, apic=True)
    ref.load(pos, vel)
    ref.initialize()
    orc = O.OracleSim((n, n, n), dx, apic=True)
    orc.load(pos, vel)
    orc.initialize()
    for f in range(frames):
        ref.update(1.0 / 30.0)
        orc.update(1.0 / 30.0)
        pr, vr = ref.particles()
        po, vo = orc.particles()
        assert bits_equal(po, pr) and bits_equal(vo, vr), f
        assert bits_equal(orc.affine(), ref.affine()), f


def test_removal_matches_reference():
    """_removeMarkerParticles (fluidsimulation.cpp:7773-7851) with both of its
    own removal rules firing: a cell crowded with 300 particles (the cap keeps
    250, counted in particle order) and 8 particles at 100 m/s, above the
    speed limit of _getMarkerParticleSpeedLimit (7723-7771).  Two frames,
    oracle against the reference bit for bit."""
    from scene_util import removal_scene
    n = 20
    dx = 1.0 / n
    pos, vel, _, _ = removal_scene(n, seed=5)
    # no gravity: the crowd stays in its cell until the cap thins it
    ref = O.ReferenceSim((n, n, n), dx, threads=1, gravity=(0.0, 0.0, 0.0))
    orc = O.OracleSim((n, n, n), dx, gravity=(0.0, 0.0, 0.0))
    for sim in (ref, orc):
        sim.load(pos, vel)
        sim.initialize()
    limit = 6 * 5 * dx * 30.0  # 6 CFL steps per frame, dt = 1/30
    for f in range(2):
        ref.update(1.0 / 30.0)
        orc.update(1.0 / 30.0)
        pr, vr = ref.particles()
        po, vo = orc.particles()
        assert bits_equal(po, pr) and bits_equal(vo, vr), f
        # both rules fired: the cap took at least 50, the speed limit the 8
        assert len(po) <= len(pos) - 58, f
        assert float(np.sqrt((vo.astype(np.float64) ** 2).sum(1)).max()) < limit, f
