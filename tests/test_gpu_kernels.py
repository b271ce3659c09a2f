"""Kernel-level parity of the HIP engine against the CPU oracle.

Every test feeds identical inputs to one kernel group through the C ABI
(ffengine.kernels -> FFAMD_* in libffengine.so) and to the oracle
restatement, then compares:
  * integer outputs (sort keys, counts, offsets, permutation, validity flags,
    pressure rows, neighbour lists): bit-exact;
  * float/double outputs whose operation order the kernel keeps (solid SDF,
    liquid SDF, P2G, extrapolation, face weights, matrix, MIC(0) factor,
    sampling): bit-exact;
  * the PCG solution, whose dot products are tree reductions on the GPU and
    sequential sums in the reference: relative tolerance stated per test.
"""
import sys

import numpy as np
import pytest

import oracle_lib as O
from ffengine import kernels as K
from ffengine import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


def same_bits(a, b):
    return np.array_equal(bits(a), bits(b))


def scene(n, seed=3, vel_scale=0.4):
    dx = 1.0 / n
    pos = scenes.dam_break_particles(n, seed=seed)
    rng = np.random.default_rng(seed + 100)
    vel = (rng.standard_normal(pos.shape) * vel_scale).astype(np.float32)
    return (n, n, n), dx, pos, vel


CASES = [24, 33]


@pytest.mark.parametrize("n", CASES)
def test_cell_sort_bitexact(n):
    n3, dx, pos, _ = scene(n)
    say(f"cell_sort n={n} P={len(pos)}")
    got = K.cell_sort(n3, dx, pos)
    want = O.cell_sort(n3, dx, pos)
    for g, w, name in zip(got, want, ("keys", "counts", "offsets", "perm")):
        assert np.array_equal(g, w), name
    # independent check: numpy's stable sort of the same keys
    assert np.array_equal(got[3], np.argsort(got[0], kind="stable"))


def test_cell_sort_edge_cases():
    n3, dx = (10, 12, 14), 0.1
    say("cell_sort empty and clustered inputs")
    keys, counts, offsets, perm = K.cell_sort(n3, dx, np.zeros((0, 3), np.float32))
    assert len(keys) == 0 and counts.sum() == 0 and offsets[-1] == 0
    # 300 particles in one cell, shuffled index order inside the cell must come out ascending
    rng = np.random.default_rng(7)
    pos = (np.array([0.55, 0.65, 0.75]) + rng.random((300, 3)) * 0.09).astype(np.float32)
    got = K.cell_sort(n3, dx, pos)
    want = O.cell_sort(n3, dx, pos)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


@pytest.mark.parametrize("dims", [(24, 24, 24), (33, 33, 33), (20, 28, 36)])
def test_solid_box_phi_bitexact(dims):
    dx = 1.0 / max(dims)
    say(f"solid_box_phi dims={dims}")
    assert same_bits(K.solid_box_phi(dims, dx), O.solid_box_phi(dims, dx))


@pytest.mark.parametrize("n", CASES)
def test_liquid_sdf_bitexact(n):
    n3, dx, pos, _ = scene(n)
    r = 0.5 * dx * np.sqrt(3.0)
    say(f"liquid_sdf n={n}")
    got = K.liquid_sdf(n3, dx, pos, r)
    want = O.liquid_sdf(n3, dx, pos, r)
    assert same_bits(got, want)
    # cell labels (liquid / air) are integer outputs of the SDF
    assert np.array_equal(got < 0, want < 0)


def test_liquid_sdf_block_boundaries():
    # particles placed exactly on and next to 10-cell block faces exercise the
    # block-membership rule
    n = 30
    dx = 1.0 / n
    g = np.array([9.0, 10.0, 10.0 + 1e-6, 19.999, 20.0], np.float64)
    P = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T * dx
    P = P.astype(np.float32)
    r = 0.5 * dx * np.sqrt(3.0)
    say("liquid_sdf block boundaries")
    assert same_bits(K.liquid_sdf((n, n, n), dx, P, r), O.liquid_sdf((n, n, n), dx, P, r))


@pytest.mark.parametrize("n", CASES)
def test_p2g_bitexact(n):
    n3, dx, pos, vel = scene(n)
    r = 0.5 * dx * np.sqrt(3.0)
    say(f"p2g n={n}")
    (gu, gv, gw), (gvu, gvv, gvw) = K.p2g(n3, dx, pos, vel, r)
    (ou, ov, ow), (ovu, ovv, ovw) = O.p2g(n3, dx, pos, vel, r)
    for g, o, name in ((gu, ou, "u"), (gv, ov, "v"), (gw, ow, "w")):
        assert same_bits(g, o), name
    for g, o in ((gvu, ovu), (gvv, ovv), (gvw, ovw)):
        assert np.array_equal(g, o)


@pytest.mark.parametrize("per_cell", [3, 40])
def test_p2g_dense_cells_bitexact(per_cell):
    """P2G on a non-cubic grid with a dense cluster: at 40 particles per cell
    the tiles over the cluster hold more particles than their LDS stage and
    read the sorted arrays instead; both paths must give the oracle's bits."""
    n3 = (24, 17, 13)
    dx = 1.0 / 24
    rng = np.random.default_rng(per_cell)
    lo, hi = np.array([4, 3, 2]), np.array([13, 11, 9])
    cells = np.stack(np.meshgrid(*[np.arange(a, b) for a, b in zip(lo, hi)], indexing="ij"), -1).reshape(-1, 3)
    cluster = (np.repeat(cells, per_cell, axis=0) + rng.random((len(cells) * per_cell, 3))) * dx
    sparse = rng.random((500, 3)) * np.array(n3) * dx
    pos = np.concatenate([cluster, sparse]).astype(np.float32)
    vel = (rng.standard_normal(pos.shape) * 0.4).astype(np.float32)
    r = 0.5 * dx * np.sqrt(3.0)
    say(f"p2g dense per_cell={per_cell} P={len(pos)}")
    (gu, gv, gw), (gvu, gvv, gvw) = K.p2g(n3, dx, pos, vel, r)
    (ou, ov, ow), (ovu, ovv, ovw) = O.p2g(n3, dx, pos, vel, r)
    for g, o, name in ((gu, ou, "u"), (gv, ov, "v"), (gw, ow, "w")):
        assert same_bits(g, o), name
    for g, o in ((gvu, ovu), (gvv, ovv), (gvw, ovw)):
        assert np.array_equal(g, o)


@pytest.mark.parametrize("n", CASES)
def test_extrapolate_bitexact(n):
    n3, dx, pos, vel = scene(n)
    r = 0.5 * dx * np.sqrt(3.0)
    (ou, ov, ow), (ovu, ovv, ovw) = O.p2g(n3, dx, pos, vel, r)
    for f, vld, name in ((ou, ovu, "u"), (ov, ovv, "v"), (ow, ovw, "w")):
        say(f"extrapolate n={n} {name}")
        assert same_bits(K.extrapolate(f, vld, 12), O.extrapolate(f, vld, 12)), name


def _solids(n, dx):
    box = O.solid_box_phi((n, n, n), dx)
    sph = scenes.sphere_phi_nodes(n, center=(0.6, 0.3, 0.5), radius=0.2)
    return scenes.union_phi(box, sph)


@pytest.mark.parametrize("n", CASES)
def test_face_weights_bitexact(n):
    dx = 1.0 / n
    phi = _solids(n, dx)
    say(f"face_weights n={n}")
    for g, o in zip(K.face_weights((n, n, n), phi), O.face_weights((n, n, n), phi)):
        assert same_bits(g, o)


def _pressure_inputs(n, seed=3):
    n3, dx, pos, vel = scene(n, seed=seed)
    r = 0.5 * dx * np.sqrt(3.0)
    solid = _solids(n, dx)
    phi = O.liquid_sdf_postprocess(n3, dx, solid, O.liquid_sdf(n3, dx, pos, r))
    (u, v, w), valid = O.p2g(n3, dx, pos, vel, r)
    u, v, w = (O.extrapolate(f, vd, 12) for f, vd in zip((u, v, w), valid))
    dt = 1.0 / 120.0
    v = (v + np.float32(-9.81 * dt)).astype(np.float32)
    weights = O.face_weights(n3, solid)
    return n3, dx, dt, (u, v, w), weights, phi


@pytest.mark.parametrize("n", CASES)
def test_pressure_system_bitexact(n):
    n3, dx, dt, uvw, weights, phi = _pressure_inputs(n)
    say(f"pressure_system n={n}")
    g = K.pressure_system(n3, dx, dt, uvw, weights, phi)
    o = O.pressure_system(n3, dx, dt, uvw, weights, phi)
    assert len(g["rows"]) == len(o["rows"]) > 0
    assert np.array_equal(g["rows"], o["rows"])
    assert np.array_equal(g["nbr"], o["nbr"])
    for key in ("rhs", "diag", "ox", "oy", "oz", "invdiag"):
        assert same_bits(g[key], o[key]), key


def _mic_sweeps(sysm, n3, r, dtype):
    """Forward and backward MIC(0) sweeps vectorised over i+j+k hyperplanes in
    `dtype`; numpy rounds every product and difference, as the kernels do."""
    ni, nj, _ = n3
    rows, nb = sysm["rows"], sysm["nbr"]
    null = sysm["diag"] < 1e-9
    inv64 = sysm["invdiag"]
    L = [np.where(null, sysm[k], sysm[k] * inv64).astype(dtype) for k in ("ox", "oy", "oz")]
    inv = inv64.astype(dtype)
    r = np.asarray(r).astype(dtype)
    z = np.zeros(len(rows), dtype)
    lev = rows % ni + (rows // ni) % nj + rows // (ni * nj)
    groups = [np.nonzero(lev == q)[0] for q in range(lev.max() + 1)]
    for g in groups:
        y = r[g].copy()
        for slot, lk in ((4, L[2]), (2, L[1]), (0, L[0])):
            k = nb[g, slot]
            h = k >= 0
            y[h] = y[h] - lk[k[h]] * z[k[h]]
        z[g] = y * inv[g]
    for g in reversed(groups):
        x = z[g].copy()
        for slot, li in ((1, L[0]), (3, L[1]), (5, L[2])):
            k = nb[g, slot]
            h = k >= 0
            x[h] = x[h] - li[g[h]] * z[k[h]]
        z[g] = x * inv[g]
    return z


def _check_mic0_apply(n, schedule):
    n3, dx, dt, uvw, weights, phi = _pressure_inputs(n)
    sysm = O.pressure_system(n3, dx, dt, uvw, weights, phi)
    m = len(sysm["rows"])
    r = np.random.default_rng(n).standard_normal(m)
    say(f"mic0_apply fp64 n={n} rows={m} schedule={schedule}")
    g64 = K.mic0_apply(n3, dx, dt, uvw, weights, phi, r, precision=0, schedule=schedule)
    o64 = O.apply_mic0(sysm, r)
    assert same_bits(g64, o64)
    assert same_bits(_mic_sweeps(sysm, n3, r, np.float64), o64)
    say(f"mic0_apply fp32 n={n} schedule={schedule}")
    g32 = K.mic0_apply(n3, dx, dt, uvw, weights, phi, r, precision=1, schedule=schedule)
    assert same_bits(g32.astype(np.float32), _mic_sweeps(sysm, n3, r, np.float32))


@pytest.mark.parametrize("schedule", [4, 2, 3, 0], ids=["line-pipeline", "bands", "bands-device-check",
                                                     "hyperplanes"])
@pytest.mark.parametrize("n", CASES + [70])
def test_mic0_apply_bitexact(n, schedule):
    """The preconditioner application as a line pipeline (pencils of lines
    along i, one launch per sweep), in bands of hyperplanes (also with the tile
    check on the device, the path of grids whose tiles outnumber a launch's
    tile mask) and one launch per hyperplane, against the oracle's serial
    sweeps (FP64) and a hyperplane-vectorised FP32 restatement (mixed
    mode)."""
    _check_mic0_apply(n, schedule=schedule)


@pytest.mark.skip(reason="tile-wavefront MIC(0) kernel switched off: under test it never finished (the compiled "
                         "ticket loop let all lanes but the ticket-taker re-run the previous tile); it stays off "
                         "the GPU until a version with a verified uniform loop has run")
@pytest.mark.parametrize("n", CASES + [70])
def test_mic0_apply_tiles_bitexact(n):
    _check_mic0_apply(n, schedule=1)


@pytest.mark.parametrize("n", CASES)
def test_pressure_solve_fp64(n):
    n3, dx, dt, uvw, weights, phi = _pressure_inputs(n)
    say(f"pressure_solve fp64 n={n}")
    gp, guvw, gvalid, git, gerr, gok = K.pressure_solve(n3, dx, dt, uvw, weights, phi, tol=1e-9, mode=0)
    op, oit, oerr, ook, m = O.pressure_solve(n3, dx, dt, uvw, weights, phi, tol=1e-9)
    assert gok and ook
    # same preconditioner and step rounding: the iteration count may move by
    # one only through the tree-ordered FP64 dot products
    assert abs(git - oit) <= 1
    scale = np.abs(op).max()
    # tolerance: the PCG stop test leaves a residual of 1e-9 |b|_inf; the two
    # solutions differ by round-off of the dot products, far below that
    assert np.abs(gp - op).max() <= 1e-6 * scale
    # the projection uses the float pressure: same bits where the pressure bits agree
    (ou, ov, ow), ovalid = O.apply_pressure(n3, dx, dt, op, phi, weights, uvw)
    for g, o in zip(gvalid, ovalid):
        assert np.array_equal(g, o)
    for g, o in zip(guvw, (ou, ov, ow)):
        assert np.abs(g - o).max() <= 1e-5 * max(np.abs(o).max(), 1e-6)


@pytest.mark.parametrize("mode", [1, 2, 3], ids=["restarted", "reliable", "fp32-preconditioner"])
@pytest.mark.parametrize("n", CASES)
def test_pressure_solve_mixed(n, mode):
    n3, dx, dt, uvw, weights, phi = _pressure_inputs(n)
    say(f"pressure_solve mixed mode={mode} n={n}")
    gp, _, _, git, gerr, gok = K.pressure_solve(n3, dx, dt, uvw, weights, phi, tol=1e-9, mode=mode)
    op, oit, oerr, ook, m = O.pressure_solve(n3, dx, dt, uvw, weights, phi, tol=1e-9)
    assert gok
    sysm = O.pressure_system(n3, dx, dt, uvw, weights, phi)
    # the FP64 true residual of the returned solution meets the reference's test
    # (up to the float rounding of the stored pressure)
    assert gerr <= 1e-9 * np.abs(sysm["rhs"]).max() * (1 + 1e-6)
    assert np.abs(gp - op).max() <= 1e-5 * np.abs(op).max()


def _divergence_max(n3, dx, dt, uvw, weights, phi):
    """max |negative divergence| over the pressure rows of a face field: the
    rhs of the system it would assemble (pressuresolver.cpp:624-713)."""
    sysm = O.pressure_system(n3, dx, dt, uvw, weights, phi)
    return float(np.abs(sysm["rhs"]).max()) if len(sysm["rhs"]) else 0.0


@pytest.mark.parametrize("mode", [0, 1, 3], ids=["fp64", "refined", "fp32-preconditioner"])
def test_precision_slab(mode):
    """The precision configuration at 40^3: a liquid slab over 40% of the
    height with a smoothed-noise velocity field (0.5 m/s, 8-cell correlation).
    Pressure against the FP64 oracle in relative L2, and the largest
    divergence left after the projection.  Tolerance: both solves meet
    |r|inf <= 1e-8 |b|inf, so their pressures differ by A^-1 (r1 - r2),
    orders of magnitude below 1e-5 in relative L2 for this Poisson matrix;
    the projection applies the pressure rounded to float, so the divergence
    left over is bounded by the float rounding of the velocity update, the
    same for the oracle and the GPU (factor 2)."""
    n = 40
    n3, dx, dt = (n, n, n), 1.0 / n, 1.0 / 120.0
    phi, uvw = scenes.precision_slab(n, seed=7)
    solid = O.solid_box_phi(n3, dx)
    weights = O.face_weights(n3, solid)
    say(f"precision slab mode={mode}")
    gp, guvw, _, git, gerr, gok = K.pressure_solve(n3, dx, dt, uvw, weights, phi, tol=1e-8, mode=mode)
    op, oit, oerr, ook, m = O.pressure_solve(n3, dx, dt, uvw, weights, phi, tol=1e-8)
    assert gok and ook and m > 0
    l2 = float(np.linalg.norm((gp - op).ravel())) / float(np.linalg.norm(op.ravel()))
    (ou, ov, ow), _ = O.apply_pressure(n3, dx, dt, op, phi, weights, uvw)
    div0 = _divergence_max(n3, dx, dt, uvw, weights, phi)
    gdiv = _divergence_max(n3, dx, dt, guvw, weights, phi)
    odiv = _divergence_max(n3, dx, dt, (ou, ov, ow), weights, phi)
    say(f"  rows={m} iters gpu={git} oracle={oit} rel L2(p)={l2:.3g} div before={div0:.3g} "
        f"after gpu={gdiv:.3g} oracle={odiv:.3g}")
    assert l2 <= 1e-5
    assert gdiv <= 2.0 * odiv + 1e-12 * div0


@pytest.mark.parametrize("n", CASES)
def test_sample_velocity(n):
    n3, dx, dt, uvw, weights, phi = _pressure_inputs(n)
    rng = np.random.default_rng(11)
    q = (rng.random((4000, 3)) * np.array(n3) * dx * 1.02 - 0.01 * n * dx).astype(np.float32)
    say(f"sample_velocity n={n}")
    g = K.sample_velocity(n3, dx, uvw, q)
    o = O.sample_velocity(n3, dx, uvw, q)
    assert np.array_equal(g, o)  # value equality (+0 == -0)
