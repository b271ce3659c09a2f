"""Scenes shared by the CPU (oracle against the reference build) and GPU
(engine against the oracle) tests of one behaviour."""
import numpy as np


def removal_scene(n, seed=0):
    """A resting block of liquid (8 jittered particles per cell, clear of the
    walls) plus the two inputs that make _removeMarkerParticles
    (fluidsimulation.cpp:7773-7851) remove particles of its own accord:

    - 300 particles crowded into one cell of the block (the cap keeps 250,
      counting in particle order, 7833-7838);
    - 8 particles in the air above the block moving at 100 m/s, above the
      speed limit _getMarkerParticleSpeedLimit derives (7723-7771): the
      histogram's top bin holds more than its removal budget, and the 8 sit
      within 10% of the fastest speed, so the outlier rule does not lower
      the limit below 6 CFL steps per frame.

    Returns (positions, velocities, crowd cell, the 8 fast particles' slice)."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / n
    lo, hi = (3, 3, 3), (n // 2, n // 2, n - 3)
    kk, jj, ii = np.meshgrid(*(np.arange(lo[a], hi[a]) for a in (2, 1, 0)), indexing="ij")
    base = np.stack([ii, jj, kk], -1).reshape(-1, 1, 3).astype(np.float64)
    block = ((base + rng.random((base.shape[0], 8, 3))) * dx).reshape(-1, 3)
    cell = np.array([lo[0] + 2, lo[1] + 2, lo[2] + 3])
    crowd = (cell + 0.05 + 0.9 * rng.random((300, 3))) * dx
    fast = (np.array([n // 2 + 2, n // 2 + 3, n // 2]) + 0.5 + 0.8 * (rng.random((8, 3)) - 0.5)) * dx
    pos = np.concatenate([block, crowd, fast]).astype(np.float32)
    vel = np.zeros_like(pos)
    dirs = np.array([0.6, 0.8, 0.0]) + 0.01 * rng.standard_normal((8, 3))
    vel[-8:] = (100.0 * dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    return pos, vel, tuple(int(c) for c in cell), slice(len(pos) - 8, len(pos))


def ball_particles(n, seed, center=(0.5, 0.55, 0.5), radius=0.22):
    """8 jittered particles per cell in the cells whose centre lies in the
    ball, float32 (P, 3)."""
    from ffengine import scenes
    pos = scenes.dam_break_particles(n, seed=seed, fraction=(1.0, 1.0, 1.0))
    dx = 1.0 / n
    cell = (np.floor(pos / dx) + 0.5) * dx
    keep = np.linalg.norm(cell - np.array(center), axis=1) < radius
    return np.ascontiguousarray(pos[keep])


def differentiation_scene(n, seed=0):
    """The differentiation configuration's scene at n^3: the dam break block
    [0, 0.375) x [0, 0.625) x [0, 1), 8 particles per cell, keeping those whose
    cell lies outside the walls (all 8 nodal solid values positive); the
    initial MAC field at rest; the target's initial field a seeded smoothed
    noise of 0.2 m/s.  Returns (positions, initial field, target field)."""
    import oracle_lib as O
    from ffengine import scenes
    dx = 1.0 / n
    pos = scenes.dam_break_particles(n, seed=seed)
    solid = O.solid_box_phi((n, n, n), dx).reshape(n + 1, n + 1, n + 1)
    c = np.floor(pos / dx).astype(int)
    free = np.ones(len(pos), bool)
    for o in range(8):
        a, b, e = o & 1, (o >> 1) & 1, o >> 2
        free &= solid[c[:, 2] + e, c[:, 1] + b, c[:, 0] + a] > 0
    shapes = [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)]
    init = [np.zeros(s, np.float32) for s in shapes]
    target_init = scenes.noise_mac_field(n, seed=seed + 1, amplitude=0.2, correlation=8.0)
    return np.ascontiguousarray(pos[free]), init, target_init


def flow_statistics(pos, vel, n, blocks=4):
    """Per-frame flow observables of a particle state, for comparing runs
    whose particles diverge individually (reordered sums amplified by the
    flow) but whose flow must agree: the particle count, the centroid and
    the mean squared speed (float64 means, in cells and (cells/s)^2), and
    the fraction of particles in each of blocks^3 equal sub-boxes of the
    domain (k, j, i order)."""
    p = np.asarray(pos, np.float64).reshape(-1, 3) * n
    v = np.asarray(vel, np.float64).reshape(-1, 3) * n
    b = np.clip(np.floor(p * blocks / n).astype(np.int64), 0, blocks - 1)
    occ = np.bincount((b[:, 2] * blocks + b[:, 1]) * blocks + b[:, 0], minlength=blocks ** 3)
    return {"count": int(len(p)), "centroid": [float(c) for c in p.mean(axis=0)],
            "mean_sq_speed": float((v * v).sum(axis=1).mean()),
            "occupancy": [float(o) for o in occ / max(len(p), 1)]}


def _moving(center, radius, sub, d, **kw):
    from ffengine import scenes
    v, t = scenes.sphere_mesh(center, radius, sub)
    d = np.asarray(d, np.float32)
    return dict(prev=v - d, cur=v, next=v + d, tris=t, **kw)


def _static(center, radius, sub, **kw):
    from ffengine import scenes
    v, t = scenes.sphere_mesh(center, radius, sub)
    return dict(prev=v, cur=v, next=v, tris=t, animated=False, **kw)


def _two_islands(d):
    from ffengine import scenes
    a, ta = scenes.sphere_mesh((0.3, 0.3, 0.3), 0.12, 1)
    b, tb = scenes.sphere_mesh((0.05, 0.7, 0.6), 0.15, 1)  # crosses the x = 0 border
    v = np.concatenate([a, b])
    t = np.concatenate([ta, tb + len(a)])
    d = np.asarray(d, np.float32)
    return dict(prev=v - d, cur=v, next=v + 2 * d, tris=t)


def moving_obstacle_cases():
    """Obstacle sets of the solid update with moving obstacles: dicts of
    prev, cur, next vertices, tris and flags (animated, inversed,
    velocity_scale).  An animated sphere beside a static one; a scaled
    animated sphere inside an inversed animated container; one animated mesh
    of two islands, one crossing the domain's border."""
    return {
        "animated-and-static": [_moving((0.45, 0.35, 0.5), 0.15, 2, (0.02, 0.01, -0.015)),
                                _static((0.7, 0.6, 0.4), 0.1, 1)],
        "scaled-and-inversed": [_moving((0.5, 0.4, 0.5), 0.12, 2, (0.0, -0.03, 0.01), velocity_scale=1.5),
                                _moving((0.5, 0.5, 0.5), 0.6, 2, (0.01, 0.0, 0.0), inversed=True)],
        "islands-across-border": [_two_islands((0.015, -0.01, 0.02))],
    }


# frame timings (frame dt, time remaining in the frame after the substep,
# substep length) of the solid-update cases: progress 0, inside the frame, and
# clamped at 0
SOLID_UPDATE_TIMINGS = [(1 / 30, 1 / 30, 1 / 120), (1 / 30, 0.5 / 30, 1 / 240), (1 / 24, 1 / 96, 1 / 96)]


def frame_progress(frame_dt, remaining, step):
    """1 - (float)(remaining + step) / (float)frame_dt, in float."""
    return float(np.float32(1.0) - np.float32(remaining + step) / np.float32(frame_dt))


def viscosity_scene(n, visc, seed=1, obstacle=True, liquid_seed=0):
    """Inputs of one viscosity solve on an n^3 grid (cell size 1/n, frame
    step 1/30): the dam-break block's liquid SDF post-processed against the
    solid, the domain box's nodal SDF (united with a sphere obstacle when
    obstacle is set), a constant viscosity on the (n+1)^3 nodes and random
    face velocities.  Returns (n3, dx, dt, liquid_phi, solid_nodes,
    viscosity, (u, v, w))."""
    import oracle_lib as O
    from ffengine import scenes
    dx = 1.0 / n
    n3 = (n, n, n)
    solid = O.solid_box_phi(n3, dx)
    if obstacle:
        solid = scenes.union_phi(solid, scenes.sphere_phi_nodes(n, center=(0.6, 0.25, 0.5), radius=0.15))
    pos = scenes.dam_break_particles(n, seed=liquid_seed)
    phi = O.liquid_sdf_postprocess(n3, dx, solid, O.liquid_sdf(n3, dx, pos, 0.5 * dx * np.sqrt(3.0)))
    rng = np.random.default_rng(seed)
    uvw = tuple(rng.standard_normal(s).astype(np.float32) for s in [(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)])
    vg = np.full((n + 1, n + 1, n + 1), visc, np.float32)
    return n3, dx, 1.0 / 30.0, phi, solid, vg, uvw
