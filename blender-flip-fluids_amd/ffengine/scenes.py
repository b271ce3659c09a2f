"""Synthetic scene inputs shared by the tests, the smoke check and the benchmark.

dam_break_particles() seeds the plain dam break: 8 particles per fluid cell,
one per half-cell sub-cube, each jittered uniformly inside its sub-cube from a
seeded numpy generator (the stratified layout of
FluidSimulation::_addMarkerParticlesToCell, fluidsimulation.cpp:7994-8004,
with the jitter made reproducible).  Velocities are zero.
"""
import numpy as np


def dam_break_particles(n, seed=0, fraction=(0.375, 0.625, 1.0), ppc_jitter=True):
    """Positions (P, 3) float32 for an n^3 grid of cell size 1/n.

    The fluid block fills cells [0, fraction_x*n) x [0, fraction_y*n) x
    [0, fraction_z*n); at n = 64 that is [0,24) x [0,40) x [0,64).
    """
    dx = 1.0 / n
    ci, cj, ck = int(round(fraction[0] * n)), int(round(fraction[1] * n)), int(round(fraction[2] * n))
    rng = np.random.default_rng(seed)
    k, j, i = np.meshgrid(np.arange(ck), np.arange(cj), np.arange(ci), indexing="ij")
    base = np.stack([i, j, k], -1).reshape(-1, 1, 3).astype(np.float64)
    sub = np.array([[a, b, c] for c in (0, 1) for b in (0, 1) for a in (0, 1)], np.float64)[None] * 0.5
    jit = rng.random((base.shape[0], 8, 3)) * 0.5 if ppc_jitter else np.full((base.shape[0], 8, 3), 0.25)
    return ((base + sub + jit) * dx).reshape(-1, 3).astype(np.float32)


def sphere_phi_nodes(n, center=(0.6, 0.2, 0.5), radius=0.15):
    """Nodal signed distance (negative inside) of a sphere on an n^3 grid of
    cell size 1/n, shaped (n+1, n+1, n+1) as (k, j, i), float32."""
    dx = 1.0 / n
    x = np.arange(n + 1, dtype=np.float64) * dx
    phi = np.empty((n + 1, n + 1, n + 1), np.float32)
    for kk in range(n + 1):
        d = np.sqrt((x[None, :] - center[0]) ** 2 + (x[:, None] - center[1]) ** 2 + (x[kk] - center[2]) ** 2)
        phi[kk] = (d - radius).astype(np.float32)
    return phi


def union_phi(*phis):
    """Union of solids: the pointwise minimum of their signed distances."""
    out = np.array(phis[0], np.float32, copy=True)
    for p in phis[1:]:
        np.minimum(out, p, out=out)
    return out
