"""The oracle's viscosity solve (oracle/fforacle.cpp viscosityVolumes,
viscositySystem, viscPCG: ViscositySolver, viscositysolver.cpp, and
PCGSolver<float>, pcgsolver.h) pinned two ways.

Against the reference built from its sources (oracle/_ref), bit for bit:
- stage by stage through the reference's own ViscositySolver: the face
  states, the seven control-volume grids, the face-to-row table, the CSR
  matrix and the right-hand side, then the whole solve (iterations, error
  and the velocity field), at three viscosities with and without an
  obstacle;
- whole frames of FluidSimulation_update with FluidSimulation_set_viscosity:
  particle positions and velocities and the frame's viscosity statistics.

Against what the mathematics fixes, with no reference at all:
- a uniform velocity field solves the system exactly (every row of the
  viscous operator sums to its volume, its solid terms included), so the
  solve returns it on every matrix face and zeroes every other face;
- the matrix is symmetric to float rounding and positive definite, and the
  PCG solution meets its stopping rule in float64 arithmetic;
- for a planar liquid surface the centre volumes are the cells' exact
  liquid fractions (trilinear interpolation of a linear SDF is exact and the
  subcell volume fraction is exact for a plane);
- zero viscosity leaves a diagonal system of the face volumes: the
  solution is the input on the matrix faces with volume, 0 off the matrix.
"""
import numpy as np
import pytest

import oracle_lib as O
from ffengine import scenes
from scene_util import viscosity_scene

needs_ref = pytest.mark.skipif(not O.reference_available(), reason="reference build (oracle/_ref) not present")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _same_system(a, b):
    assert a["m"] == b["m"]
    for x, y in zip(a["states"], b["states"]):
        assert np.array_equal(x, y)
    for k in O.VISC_VOLUMES:
        assert np.array_equal(bits(a["volumes"][k]), bits(b["volumes"][k])), k
    for k in ("table", "rowstart", "cols"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(bits(a["vals"]), bits(b["vals"]))
    assert np.array_equal(bits(a["rhs"]), bits(b["rhs"]))


@needs_ref
@pytest.mark.parametrize("n,visc,obstacle", [(16, 0.1, True), (20, 0.005, False), (24, 2.0, True)])
def test_stages_match_reference(n, visc, obstacle):
    n3, dx, dt, phi, solid, vg, uvw = viscosity_scene(n, visc, obstacle=obstacle)
    mine = O.viscosity_system(n3, dx, dt, phi, solid, vg, uvw)
    _same_system(mine, O.ref_viscosity_system(n3, dx, dt, phi, solid, vg, uvw))
    assert mine["m"] > 1000
    got, it, err, ok, m = O.viscosity_solve(n3, dx, dt, phi, solid, vg, uvw)
    want, rit, rerr, rok = O.ref_viscosity_solve(n3, dx, dt, phi, solid, vg, uvw)
    assert (it, ok) == (rit, rok) and np.float32(err) == np.float32(rerr)
    assert ok and it >= 3
    for a, b in zip(got, want):
        assert np.array_equal(bits(a), bits(b))


@needs_ref
def test_variable_viscosity_grid_matches_reference():
    """A non-constant viscosity grid exercises the neighbour averages of
    every row type (a transposed average or a wrong node fails it)."""
    n3, dx, dt, phi, solid, vg, uvw = viscosity_scene(18, 0.1)
    rng = np.random.default_rng(7)
    vg = (0.02 + rng.random(vg.shape)).astype(np.float32)
    _same_system(O.viscosity_system(n3, dx, dt, phi, solid, vg, uvw),
                 O.ref_viscosity_system(n3, dx, dt, phi, solid, vg, uvw))
    got = O.viscosity_solve(n3, dx, dt, phi, solid, vg, uvw)
    want = O.ref_viscosity_solve(n3, dx, dt, phi, solid, vg, uvw)
    assert got[1:4] == want[1:4]
    for a, b in zip(got[0], want[0]):
        assert np.array_equal(bits(a), bits(b))


@needs_ref
@pytest.mark.parametrize("visc,obstacle", [(0.5, False), (3.0, True)])
def test_frames_match_reference(visc, obstacle):
    n = 16
    dx = 1.0 / n
    pos = scenes.dam_break_particles(n, seed=0)
    ref = O.ReferenceSim((n, n, n), dx, threads=1)
    orc = O.OracleSim((n, n, n), dx)
    if obstacle:
        v, t = scenes.sphere_mesh((0.6, 0.25, 0.5), 0.15, 2)
        ref.add_mesh_obstacle((n, n, n), dx, v, t)
        orc.add_moving_obstacle(v, v, v, t, animated=False)
    for s in (ref, orc):
        s.load(pos)
        s.set_viscosity(visc)
        s.initialize()
    for f in range(3):
        ref.update(1.0 / 30.0)
        orc.update(1.0 / 30.0)
        rp, rv = ref.particles()
        op, ov = orc.particles()
        assert len(rp) == len(op), f
        assert np.array_equal(bits(rp), bits(op)), f
        assert np.array_equal(bits(rv), bits(ov)), f
        rs, os_ = ref.viscosity_stats(), orc.viscosity_stats()
        assert rs[:2] == os_[:2] and np.float32(rs[2]) == np.float32(os_[2]), f
        assert os_[0] and os_[1] > 3


def _full_liquid(n, obstacle):
    n3, dx, dt, _, solid, vg, _ = viscosity_scene(n, 0.3, obstacle=obstacle)
    phi = np.full((n, n, n), -0.5 * dx, np.float32)  # liquid everywhere
    return n3, dx, dt, phi, solid, vg


@pytest.mark.parametrize("obstacle", [False, True])
def test_uniform_field_is_the_exact_solution(obstacle):
    n = 14
    n3, dx, dt, phi, solid, vg = _full_liquid(n, obstacle)
    c = (0.7, -1.3, 0.4)
    uvw = tuple(np.full(s, cc, np.float32) for s, cc in zip([(n, n, n + 1), (n, n + 1, n), (n + 1, n, n)], c))
    S = O.viscosity_system(n3, dx, dt, phi, solid, vg, uvw)
    out, it, err, ok, m = O.viscosity_solve(n3, dx, dt, phi, solid, vg, uvw, tol=1e-6)
    assert ok and m == S["m"] > 0
    nU = (n + 1) * n * n
    nV = n * (n + 1) * n
    tables = (S["table"][:nU], S["table"][nU:nU + nV], S["table"][nU + nV:])
    for a, cc, tab in zip(out, c, tables):
        flat = a.reshape(-1)
        inm = tab >= 0
        assert np.all(flat[~inm] == 0.0)
        assert np.abs(flat[inm] - cc).max() < 1e-4 * abs(cc)


def _dense(S):
    m = S["m"]
    A = np.zeros((m, m))
    for r in range(m):
        lo, hi = S["rowstart"][r], S["rowstart"][r + 1]
        A[r, S["cols"][lo:hi]] = S["vals"][lo:hi]
    return A


def test_matrix_is_symmetric_positive_definite_and_pcg_meets_its_rule():
    n3, dx, dt, phi, solid, vg, uvw = viscosity_scene(12, 0.8)
    S = O.viscosity_system(n3, dx, dt, phi, solid, vg, uvw)
    A = _dense(S)
    assert np.abs(A - A.T).max() <= 1e-6 * np.abs(A).max()
    assert np.linalg.eigvalsh(0.5 * (A + A.T)).min() > 0
    x, it, err, ok = O.viscosity_pcg(S, tol=1e-4)
    b = S["rhs"].astype(np.float64)
    assert ok and it > 3
    assert np.abs(A @ x - b).max() <= 1.5e-4 * np.abs(b).max()


def test_centre_volumes_of_a_planar_surface():
    n = 16
    dx = 1.0 / n
    h = 0.41
    n3, _, dt, _, _, vg, uvw = viscosity_scene(n, 0.1, obstacle=False)
    solid = O.solid_box_phi(n3, dx)
    y = (np.arange(n) + 0.5) * dx
    phi = np.broadcast_to((y - h)[None, :, None], (n, n, n)).astype(np.float32)
    S = O.viscosity_system(n3, dx, dt, phi, solid, vg, uvw)
    want = np.clip((h - np.arange(n) * dx) / dx, 0.0, 1.0)
    got = S["volumes"]["center"]
    inner = got[2:n - 2, 2:n - 2, 2:n - 2]
    assert np.abs(inner - want[None, 2:n - 2, None]).max() < 1e-5
    assert np.any((inner > 0.2) & (inner < 0.8))


def test_zero_viscosity_keeps_the_field_on_faces_with_volume():
    n3, dx, dt, phi, solid, vg, uvw = viscosity_scene(14, 0.0)
    S = O.viscosity_system(n3, dx, dt, phi, solid, vg, uvw)
    out, it, err, ok, m = O.viscosity_solve(n3, dx, dt, phi, solid, vg, uvw)
    assert ok
    vols = (S["volumes"]["U"], S["volumes"]["V"], S["volumes"]["W"])
    sizes = np.cumsum([0] + [a.size for a in uvw])
    for d, (a, b, vol) in enumerate(zip(out, uvw, vols)):
        tab = S["table"][sizes[d]:sizes[d + 1]].reshape(a.shape)
        has = (vol > 1e-3) & (tab >= 0)
        assert has.sum() > 100
        assert np.abs(a[has] - b[has]).max() < 1e-3 * np.abs(b).max()
        assert np.all(a[tab < 0] == 0.0)
