"""The oracle's moving-solid terms of the pressure solve, pinned by what the
mathematics fixes (they are also pinned as part of whole frames against the
reference engine in tests/test_oracle_moving_obstacles.py):

- cell weights (oracle cellWeights: fluidsimulation.cpp:6140-6147,
  levelsetutils.cpp:213-258): for a linear SDF the piecewise-linear
  interpolation on the cube's tetrahedra is exact, so 1 - weight is the
  volume of the cube below a plane, which has the closed form
  sum over corners v of (-1)^|v| max(0, d - a.v)^3 / (6 a1 a2 a3);
  a dropped tetrahedron, a wrong vertex or a missing factor 2 fails it;
- isolated-pocket conditioning (oracle conditionSolidVelocity:
  pressuresolver.cpp:161-312): the groups are the connected components of
  the interior liquid cells joined through faces of weight >= 1e-6, found
  here with scipy's connected_components on that graph; the faces of the
  cells of a component of more than one cell with no open face to an air cell
  are zeroed, every other face keeps its value.
"""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

import oracle_lib as O


def _cube_below_plane(a, d):
    """Volume of {x in [0,1]^3 : a.x < d} (all a nonzero)."""
    a = np.asarray(a, np.float64)
    # reflect negative axes: a_i x_i = a_i - |a_i| (1 - x_i)
    d = d - a[a < 0].sum()
    a = np.abs(a)
    vol = 0.0
    for v in np.ndindex(2, 2, 2):
        t = d - float(np.dot(a, v))
        if t > 0:
            vol += (-1) ** sum(v) * t ** 3
    return min(max(vol / (6.0 * a.prod()), 0.0), 1.0)


@pytest.mark.parametrize("normal,offset", [((0.0, 0.0, 1.0), 0.37), ((0.3, -0.5, 0.81), 0.52),
                                           ((-0.7, 0.2, -0.4), -0.18)])
def test_cell_weights_of_a_plane(normal, offset):
    n = 12
    dx = 1.0 / n
    a = np.asarray(normal, np.float64)
    if np.any(a == 0):
        a = a + np.where(a == 0, 1e-4, 0.0)  # the closed form needs nonzero components
    a = a / np.linalg.norm(a)
    z, y, x = np.meshgrid(*(np.arange(n + 1) * dx,) * 3, indexing="ij")
    phi = (a[0] * x + a[1] * y + a[2] * z - offset).astype(np.float32)  # solid where phi < 0
    wc = O.cell_weights((n, n, n), phi)
    want = np.zeros((n, n, n))
    for k in range(n):
        for j in range(n):
            for i in range(n):
                # in the cell's unit frame: a.(corner + dx t) < offset
                d = (offset - a[0] * i * dx - a[1] * j * dx - a[2] * k * dx) / dx
                want[k, j, i] = 1.0 - _cube_below_plane(a, d)
    err = np.abs(wc.astype(np.float64) - want).max()
    assert err < 2e-4, err
    assert wc.min() >= 0.0 and wc.max() <= 1.0
    assert np.any((wc > 0.05) & (wc < 0.95))  # the plane does cut cells


def test_cell_weights_inside_and_outside():
    n = 8
    phi = np.full((n + 1, n + 1, n + 1), 0.3, np.float32)
    assert np.all(O.cell_weights((n, n, n), phi) == 1.0)
    assert np.all(O.cell_weights((n, n, n), -phi) == 0.0)


def _pockets_expected(n, phi, wu, wv, ww, us, vs, ws):
    eps = 1e-6
    idx = lambda i, j, k: (k * n + j) * n + i
    member = np.zeros(n ** 3, bool)
    for k in range(1, n - 1):
        for j in range(1, n - 1):
            for i in range(1, n - 1):
                member[idx(i, j, k)] = phi[k, j, i] < 0
    rows, cols = [], []
    air = np.zeros(n ** 3, bool)
    for k in range(1, n - 1):
        for j in range(1, n - 1):
            for i in range(1, n - 1):
                c = idx(i, j, k)
                faces = [(wu[k, j, i], (i - 1, j, k)), (wu[k, j, i + 1], (i + 1, j, k)),
                         (wv[k, j, i], (i, j - 1, k)), (wv[k, j + 1, i], (i, j + 1, k)),
                         (ww[k, j, i], (i, j, k - 1)), (ww[k + 1, j, i], (i, j, k + 1))]
                for wt, (a, b, d) in faces:
                    if wt >= eps and phi[d, b, a] >= 0:
                        air[c] = True
                    if member[c] and member[idx(a, b, d)] and wt >= eps:
                        rows.append(c)
                        cols.append(idx(a, b, d))
    g = coo_matrix((np.ones(len(rows)), (rows, cols)), shape=(n ** 3, n ** 3))
    _, label = connected_components(g, directed=False)
    us, vs, ws = us.copy(), vs.copy(), ws.copy()
    for lab in np.unique(label[member]):
        cells = np.nonzero(member & (label == lab))[0]
        if len(cells) <= 1 or air[cells].any():
            continue
        for c in cells:
            i, j, k = c % n, (c // n) % n, c // (n * n)
            us[k, j, i] = us[k, j, i + 1] = 0.0
            vs[k, j, i] = vs[k, j + 1, i] = 0.0
            ws[k, j, i] = ws[k + 1, j, i] = 0.0
    return us, vs, ws


def pocket_scene(n, seed):
    """Liquid with walls of closed faces: some regions sealed off from the air."""
    rng = np.random.default_rng(seed)
    phi = np.where(rng.random((n, n, n)) < 0.8, -0.5, 0.5).astype(np.float32)
    wu = np.ones((n, n, n + 1), np.float32)
    wv = np.ones((n, n + 1, n), np.float32)
    ww = np.ones((n + 1, n, n), np.float32)
    for w in (wu, wv, ww):
        r = rng.random(w.shape)
        w[r < 0.35] = 0.0
        w[(r >= 0.35) & (r < 0.4)] = 5e-7  # below the 1e-6 threshold
        w[(r >= 0.4) & (r < 0.5)] = 0.25
    # a sealed box of liquid: closed faces all round cells [3, 6)^3
    phi[3:6, 3:6, 3:6] = -0.5
    wu[3:6, 3:6, 3:7] = 1.0
    wv[3:6, 3:7, 3:6] = 1.0
    ww[3:7, 3:6, 3:6] = 1.0
    wu[3:6, 3:6, 3] = wu[3:6, 3:6, 6] = 0.0
    wv[3:6, 3, 3:6] = wv[3:6, 6, 3:6] = 0.0
    ww[3, 3:6, 3:6] = ww[6, 3:6, 3:6] = 0.0
    us = rng.standard_normal(wu.shape).astype(np.float32)
    vs = rng.standard_normal(wv.shape).astype(np.float32)
    ws = rng.standard_normal(ww.shape).astype(np.float32)
    return phi, (wu, wv, ww), (us, vs, ws)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_conditioning_zeroes_isolated_pockets(seed):
    n = 12
    phi, wts, vel = pocket_scene(n, seed)
    got = O.condition_solid_velocity((n, n, n), phi, wts, vel)
    want = _pockets_expected(n, phi, *wts, *vel)
    for a, b in zip(got, want):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.all(got[0][3:6, 3:6, 3:7] == 0.0)  # the sealed box
    assert any(np.count_nonzero(a != b) for a, b in zip(got, vel))
