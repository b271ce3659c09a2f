"""Scenes of the surface tension checks and the reference-side runs that
tests/test_oracle_surface_tension.py starts in a fresh interpreter
(oracle_lib.run_reference_isolated): each function returns a dict of arrays."""
import numpy as np

import oracle_lib as O
from ffengine import scenes
from scene_util import ball_particles


def tension_scene(n, seed=3, drops=40):
    """The dam-break block's liquid SDF plus small droplets scattered in the
    air (they make small, BAD clusters): (n3, dx, solid phi, liquid phi)."""
    dx = 1.0 / n
    n3 = (n, n, n)
    rng = np.random.default_rng(seed)
    pos = np.concatenate([scenes.dam_break_particles(n, seed=0),
                          (rng.random((drops, 3)) * 0.8 + 0.1).astype(np.float32)])
    solid = O.solid_box_phi(n3, dx)
    phi = O.liquid_sdf_postprocess(n3, dx, solid, O.liquid_sdf(n3, dx, pos, 0.5 * dx * np.sqrt(3.0)))
    return n3, dx, solid, phi


def random_field(n, seed=1):
    rng = np.random.default_rng(seed)
    return [0.3 * rng.standard_normal(s).astype(np.float32) for s in [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)]]


def ref_curvature(n, drops):
    n3, dx, solid, phi = tension_scene(n, drops=drops)
    surf, kg = O.ref_curvature_grid(n3, dx, phi)
    return {"surface": surf, "curvature": kg}


def ref_pressure(n, k):
    n3, dx, solid, phi = tension_scene(n)
    p, vel, flags, it, err, ok, cl = O.ref_pressure_solve_tension(n3, dx, 1.0 / 60.0, random_field(n),
                                                                  O.face_weights(n3, solid), phi, k)
    return {"p": p, "u": vel[0], "v": vel[1], "w": vel[2], "vu": flags[0], "vv": flags[1], "vw": flags[2],
            "it": np.array(it), "ok": np.array(ok), "cluster": cl}


def ref_frames(n, k, smooth, frames):
    dx = 1.0 / n
    ref = O.ReferenceSim((n, n, n), dx, threads=1)
    ref.load(ball_particles(n, 0))
    ref.set_surface_tension(k, smooth=smooth)
    ref.initialize()
    out = {}
    for f in range(frames):
        ref.update(1.0 / 30.0)
        out["p%d" % f], out["v%d" % f] = ref.particles()
    return out
