"""The oracle's surface tension (oracle/fforacle.cpp curvatureGrid,
tensionClusterStatus, Tension and the tension terms of buildSystem and
applyPressure) pinned two ways.

Against the reference built from its sources (oracle/_ref), bit for bit:
- the reinitialized surface phi and the extrapolated curvature grid of
  ParticleLevelSet::calculateCurvatureGrid (particlelevelset.cpp:196-292);
- the pressure solve with surface tension through the reference's own
  PressureSolver (pressure, velocities, valid flags, iterations) and its
  cluster status (pressuresolver.cpp:314-606, 624-749, 1047-1205);
- whole frames of FluidSimulation_update with FluidSimulation_set_surface_tension,
  with the smooth kernel, and with a constant strong enough for the
  substep restriction (fluidsimulation.cpp:10248-10251) to add substeps.
The reference's level set solver reads a buffer it never initializes
(levelsetsolver.cpp:97), so its result depends on what that memory held;
the reference therefore runs in a fresh interpreter whose large allocations
come from zeroed pages (oracle_lib.run_reference_isolated), where the buffer
reads 0 as the oracle assumes.

Against what the mathematics fixes, with no reference at all:
- the curvature of a sphere's exact SDF is 2/R at its surface (median
  within 2%), and a plane's SDF is kept with no curvature;
- the cluster status marks BAD exactly the cells of the connected
  same-phase regions (6-neighbours, outside the 4^3 blocks that lie wholly
  in one phase and off the border) that hold fewer than 36 cells and touch
  no such block, found here with scipy's connected components.
"""
import numpy as np
import pytest
from scipy import ndimage

import oracle_lib as O
from scene_util import ball_particles
from tension_scenes import random_field, tension_scene

needs_ref = pytest.mark.skipif(not O.reference_available(), reason="reference build (oracle/_ref) not present")


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@needs_ref
@pytest.mark.parametrize("n,drops", [(24, 0), (24, 40), (30, 40), (40, 40)])
def test_curvature_grid_matches_reference(n, drops, tmp_path):
    n3, dx, solid, phi = tension_scene(n, drops=drops)
    surf, kg = O.curvature_grid(n3, dx, phi)
    ref = O.run_reference_isolated("tension_scenes", "ref_curvature", str(tmp_path / "r.npz"), n=n, drops=drops)
    assert np.array_equal(bits(surf), bits(ref["surface"]))
    assert np.array_equal(bits(kg), bits(ref["curvature"]))
    assert np.count_nonzero(kg) > 100


@needs_ref
@pytest.mark.parametrize("k", [0.05, 2.0])
def test_pressure_with_tension_matches_reference(k, tmp_path):
    n = 24
    n3, dx, solid, phi = tension_scene(n)
    wts = O.face_weights(n3, solid)
    uvw = random_field(n)
    p, vel, flags, it, err, ok = O.pressure_solve_tension(n3, dx, 1.0 / 60.0, uvw, wts, phi, k)
    ref = O.run_reference_isolated("tension_scenes", "ref_pressure", str(tmp_path / "r.npz"), n=n, k=k)
    assert ok and bool(ref["ok"]) and it == int(ref["it"])
    assert np.array_equal(bits(p), bits(ref["p"]))
    for a, key in zip(vel, "uvw"):
        assert np.array_equal(bits(a), bits(ref[key]))
    for a, key in zip(flags, ("vu", "vv", "vw")):
        assert np.array_equal(a, ref[key])
    st = O.tension_cluster_status(n3, phi)
    assert np.array_equal(st, ref["cluster"])
    assert np.count_nonzero(st & 12) > 0  # BAD clusters present
    # the terms move the pressure
    p0 = O.pressure_solve_tension(n3, dx, 1.0 / 60.0, uvw, wts, phi, 0.0)[0]
    assert np.abs(p - p0).max() > 1e-4


@needs_ref
@pytest.mark.parametrize("n,k,smooth,frames", [(24, 0.5, False, 3), (24, 5.0, True, 3), (32, 50.0, False, 2)])
def test_frames_match_reference(n, k, smooth, frames, tmp_path):
    dx = 1.0 / n
    ref = O.run_reference_isolated("tension_scenes", "ref_frames", str(tmp_path / "r.npz"), n=n, k=k, smooth=smooth,
                                   frames=frames)
    orc = O.OracleSim((n, n, n), dx)
    orc.load(ball_particles(n, 0))
    orc.set_surface_tension(k, smooth=smooth)
    orc.initialize()
    most = 1
    for f in range(frames):
        orc.update(1.0 / 30.0)
        op, ov = orc.particles()
        assert len(op) == len(ref["p%d" % f]), f
        assert np.array_equal(bits(op), bits(ref["p%d" % f])), f
        assert np.array_equal(bits(ov), bits(ref["v%d" % f])), f
        most = max(most, orc.substeps())
    if k >= 50:
        assert most > 1  # the surface tension restriction adds substeps


def test_sphere_curvature_is_two_over_radius():
    """The mean curvature of a sphere of radius R is 2/R.  The upwind
    reinitialization moves the field by a few hundredths of a cell, which the
    second differences amplify cell by cell, so the median over the cells at
    the surface is held to 2% and the middle half to 20%."""
    n, R = 48, 0.3
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    phi = (np.sqrt((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2) - R).astype(np.float32)
    surf, kg = O.curvature_grid((n, n, n), dx, phi)
    near = np.abs(phi) < 0.5 * dx
    assert near.sum() > 1000
    q25, med, q75 = np.percentile(kg[near], [25, 50, 75])
    assert abs(med - 2.0 / R) < 0.02 * (2.0 / R), med
    assert 0.8 * (2.0 / R) < q25 and q75 < 1.2 * (2.0 / R), (q25, q75)
    assert np.abs(surf[near] - phi[near]).max() < 0.05 * dx  # an SDF stays close to itself


def test_plane_keeps_its_distance_and_has_no_curvature():
    """A level plane's SDF already has |grad phi| = 1, and with the surface
    across the middle of a 24^3 grid every cell is a solver cell: the
    reinitialization keeps it (to rounding) and the curvature vanishes."""
    n = 24
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx
    phi = np.broadcast_to((c - 0.4937)[None, :, None], (n, n, n)).astype(np.float32)
    surf, kg = O.curvature_grid((n, n, n), dx, phi)
    assert np.abs(surf - phi).max() < 1e-6
    near = np.abs(phi) < 1.5 * dx
    assert np.abs(kg[near]).max() < 1e-3 / dx


@pytest.mark.parametrize("seed", [3, 4])
def test_bad_clusters_are_the_small_isolated_regions(seed):
    n = 24
    n3, dx, solid, phi = tension_scene(n, seed=seed, drops=60)
    st = O.tension_cluster_status(n3, phi)
    border = np.zeros(n3[::-1], bool)
    border[0, :, :] = border[-1, :, :] = border[:, 0, :] = border[:, -1, :] = True
    border[:, :, 0] = border[:, :, -1] = True
    inside = phi < 0
    # cells of 4^3 blocks wholly in one phase
    preset = np.zeros_like(border)
    b = n // 4
    blk = inside[:4 * b, :4 * b, :4 * b].reshape(b, 4, b, 4, b, 4)
    allin = blk.all(axis=(1, 3, 5))
    allout = (~blk).all(axis=(1, 3, 5))
    full = np.repeat(np.repeat(np.repeat(allin | allout, 4, 0), 4, 1), 4, 2)
    preset[:4 * b, :4 * b, :4 * b] = full
    preset &= ~border
    free = ~border & ~preset
    want = np.zeros_like(border)
    six = ndimage.generate_binary_structure(3, 1)
    for phase in (inside, ~inside):
        lab, nl = ndimage.label(free & phase, structure=six)
        touch = ndimage.binary_dilation(preset & phase, structure=six) & (lab > 0)
        touched = set(np.unique(lab[touch]).tolist())
        sizes = ndimage.sum(np.ones_like(lab), lab, index=np.arange(1, nl + 1))
        for lab_id in range(1, nl + 1):
            if sizes[lab_id - 1] < 36 and lab_id not in touched:
                want |= lab == lab_id
    got = (st & 12) != 0
    assert np.array_equal(got, want)
    assert want.sum() > 0
