"""Synthetic scene inputs shared by the tests, the smoke check and the benchmark.

dam_break_particles() seeds the plain dam break: 8 particles per fluid cell,
one per half-cell sub-cube, each jittered uniformly inside its sub-cube from a
seeded numpy generator (the stratified layout of
FluidSimulation::_addMarkerParticlesToCell, fluidsimulation.cpp:7994-8004,
with the jitter made reproducible).  Velocities are zero.

precision_slab() builds the grid state of the precision configuration: a
liquid slab over the bottom 40% of the domain whose MAC velocity is seeded
Gaussian noise smoothed to an 8-cell correlation length.
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


def sphere_mesh(center=(0.6, 0.2, 0.5), radius=0.15, subdivisions=4):
    """Closed triangle mesh of a sphere: an icosahedron subdivided
    `subdivisions` times (20 * 4**subdivisions triangles) with its vertices
    projected onto the sphere.  Returns (vertices (V, 3) float32, triangles
    (T, 3) int32)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    tris = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
            (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
            (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        mid = {}

        def midpoint(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in mid:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                mid[key] = len(verts) - 1
            return mid[key]

        nxt = []
        for a, b, c in tris:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nxt += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        tris = nxt
    v = np.array(verts) * radius + np.asarray(center, np.float64)
    return v.astype(np.float32), np.array(tris, np.int32)


def union_phi(*phis):
    """Union of solids: the pointwise minimum of their signed distances."""
    out = np.array(phis[0], np.float32, copy=True)
    for p in phis[1:]:
        np.minimum(out, p, out=out)
    return out


def noise_mac_field(n, seed=0, amplitude=0.5, correlation=8.0):
    """u, v, w face arrays ((k, j, i) shapes (n, n, n+1), (n, n+1, n),
    (n+1, n, n)), float32: independent Gaussian noise per component, smoothed
    by a Gaussian filter of standard deviation correlation/2 cells and scaled
    to a standard deviation of `amplitude` (m/s)."""
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(seed)
    out = []
    for shape in ((n, n, n + 1), (n, n + 1, n), (n + 1, n, n)):
        f = gaussian_filter(rng.standard_normal(shape, dtype=np.float32), sigma=correlation / 2.0, mode="wrap")
        f *= np.float32(amplitude / max(float(f.std()), 1e-30))
        out.append(f.astype(np.float32))
    return tuple(out)


def slab_phi(n, height=0.4):
    """Liquid signed distance at cell centres ((k, j, i), float32): negative
    below y = height (the slab), the distance to its top face elsewhere."""
    dx = 1.0 / n
    y = (np.arange(n, dtype=np.float64) + 0.5) * dx
    phi = np.empty((n, n, n), np.float32)
    phi[:] = (y - height).astype(np.float32)[None, :, None]
    return phi


def precision_slab(n, seed=0, height=0.4, amplitude=0.5, correlation=8.0):
    """Grid state of the precision configuration: (liquid phi, (u, v, w))."""
    return slab_phi(n, height), noise_mac_field(n, seed, amplitude, correlation)


def splash_particles(n, seed=0, pool=0.3, center=(0.5, 0.65, 0.5), radius=0.2):
    """8 jittered particles per cell in the pool (y < pool) and in the cells
    whose centre lies inside the falling sphere of liquid."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx
    chunks = []
    for k in range(n):
        zz, yy, xx = c[k], c[:, None], c[None, :]
        inside = (yy < pool) | ((xx - center[0]) ** 2 + (yy - center[1]) ** 2 + (zz - center[2]) ** 2 < radius ** 2)
        jj, ii = np.nonzero(np.broadcast_to(inside, (n, n)))
        if len(ii) == 0:
            continue
        base = np.stack([ii, jj, np.full_like(ii, k)], -1).reshape(-1, 1, 3).astype(np.float32)
        sub = np.array([[a, b, cc] for cc in (0, 1) for b in (0, 1) for a in (0, 1)], np.float32)[None] * 0.5
        jit = rng.random((base.shape[0], 8, 3), dtype=np.float32) * 0.5
        chunks.append(((base + sub + jit) * np.float32(dx)).reshape(-1, 3))
    return np.concatenate(chunks).astype(np.float32)
