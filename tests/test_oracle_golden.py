"""The CPU oracle against the golden fixtures written from the reference engine
(tests/golden/make_golden.py).  Every comparison is bit-exact: the oracle keeps
the reference's float/double promotion points and operation order.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    b = np.ascontiguousarray(b, np.float32).reshape(-1)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module")
def st():
    with np.load(os.path.join(GOLDEN, "stages_n16.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _n3(st):
    n = int(st["n"])
    return (n, n, n), float(st["dx"])


def test_solid_box_phi(st):
    n3, dx = _n3(st)
    assert bits_equal(O.solid_box_phi(n3, dx), st["solid_box_phi"])


def test_liquid_sdf(st):
    n3, dx = _n3(st)
    raw = O.liquid_sdf(n3, dx, st["pos"], float(st["radius"]))
    assert bits_equal(raw, st["liquid_sdf"])
    n = n3[0]
    post = O.liquid_sdf_postprocess(n3, dx, st["solid_phi"].reshape(n + 1, n + 1, n + 1), raw)
    assert bits_equal(post, st["liquid_phi"])


def test_p2g_and_extrapolation(st):
    n3, dx = _n3(st)
    fields, valids = O.p2g(n3, dx, st["pos"], st["vel"], float(st["radius"]))
    for d, name in enumerate("uvw"):
        assert bits_equal(fields[d], st["p2g_" + name]), name
        assert np.array_equal(valids[d].reshape(-1), st["p2g_valid_" + name]), name
        ext = O.extrapolate(fields[d], valids[d], 12)
        assert bits_equal(ext, st["extrap_" + name]), name


def test_face_weights(st):
    n3, dx = _n3(st)
    n = n3[0]
    w = O.face_weights(n3, st["solid_phi"].reshape(n + 1, n + 1, n + 1))
    for d, name in enumerate("uvw"):
        assert bits_equal(w[d], st["weight_" + name]), name


def _shapes(n):
    return [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)]


def test_pressure_solve_and_apply(st):
    n3, dx = _n3(st)
    n = n3[0]
    dt = float(st["dt"])
    shp = _shapes(n)
    uvw = [st["pre_" + c].reshape(s) for c, s in zip("uvw", shp)]
    wts = [st["weight_" + c].reshape(s) for c, s in zip("uvw", shp)]
    phi = st["liquid_phi"].reshape(n, n, n)
    p, iters, err, ok, m = O.pressure_solve(n3, dx, dt, uvw, wts, phi, tol=float(st["tol"]))
    assert ok and m > 0
    assert iters == int(st["pressure_iterations"])
    assert np.float32(err) == st["pressure_error"]
    assert bits_equal(p, st["pressure"])
    out, valid = O.apply_pressure(n3, dx, dt, p, phi, wts, uvw)
    for d, c in enumerate("uvw"):
        assert bits_equal(out[d], st["post_" + c]), c
        assert np.array_equal(valid[d].reshape(-1), st["post_valid_" + c]), c


def test_sample_velocity(st):
    n3, dx = _n3(st)
    n = n3[0]
    uvw = [st["post_" + c].reshape(s) for c, s in zip("uvw", _shapes(n))]
    got = O.sample_velocity(n3, dx, uvw, st["sample_points"])
    assert bits_equal(got, st["sample_velocity"])


def test_engine_frames_16():
    with np.load(os.path.join(GOLDEN, "e2e_n16.npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    n = int(g["n"])
    sim = O.OracleSim((n, n, n), float(g["dx"]))
    sim.load(g["pos0"])
    sim.initialize()
    for f in range(1, 11):
        sim.update(float(g["frame_dt"]))
        if "pos_f%d" % f in g:
            p, v = sim.particles()
            assert bits_equal(p, g["pos_f%d" % f]), f
            assert bits_equal(v, g["vel_f%d" % f]), f


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _dam_break(n, seed):
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(GOLDEN, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.dam_break(n, seed)


@pytest.mark.parametrize("frames", [3, 60])
def test_engine_frames_64(frames):
    """The 64^3 dam break: checksums of every frame.  The full 60 frames take
    a few minutes of CPU and run when FF_ORACLE_LONG=1."""
    path = os.path.join(GOLDEN, "e2e_c1_n64.json")
    if not os.path.isfile(path):
        pytest.skip("e2e_c1_n64.json not generated")
    if frames > 3 and os.environ.get("FF_ORACLE_LONG") != "1":
        pytest.skip("set FF_ORACLE_LONG=1 for the 60-frame run")
    with open(path) as fh:
        rec = json.load(fh)
    n = 64
    pos = _dam_break(n, 0)
    assert _sha(pos) == rec["initial_positions_sha256"]
    sim = O.OracleSim((n, n, n), 1.0 / n)
    sim.load(pos)
    sim.initialize()
    for fr in rec["frames"][:frames]:
        sim.update(1.0 / 30.0)
        p, v = sim.particles()
        assert len(p) == fr["count"], fr["frame"]
        assert _sha(p) == fr["positions_sha256"], fr["frame"]
        assert _sha(v) == fr["velocities_sha256"], fr["frame"]
