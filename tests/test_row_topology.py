"""The line words' row lookup used by the MIC(0) line pipeline
(csrc/pencil_sweep.h rowAt, csrc/kernels_pencil.hip lineWordsKernel)
restated on the CPU and checked against the oracle's neighbour lists
(pressuresolver.cpp:118-159 row order, the 7-point stencil of
sparsematrix.h:174-215).

The lookup needs no per-row neighbour ids.  It finds:
  * the +-x neighbour rows as r -+ 1, present when rows[r -+ 1] is the adjacent
    cell;
  * the +-y/+-z neighbour rows in the line words: for each 32-cell word of a
    (j, k) line along i, the bit mask of its pressure cells and the row of its
    first pressure cell (the exclusive scan of the row mask at the word's first
    cell), row = base + popcount(bits below the cell).
Grids whose i extent is not a multiple of 32 (33, 70) exercise the partial
last word of a line.  The oracle's neighbour lists are its own assembly's,
so a wrong word index, bit or base fails here before any GPU run.
"""
import numpy as np
import pytest

import oracle_lib as O
from ffengine import scenes


def _system(n, seed):
    dx = 1.0 / n
    pos = scenes.dam_break_particles(n, seed=seed)
    n3 = (n, n, n)
    r = 0.5 * dx * np.sqrt(3.0)
    box = O.solid_box_phi(n3, dx)
    solid = scenes.union_phi(box, scenes.sphere_phi_nodes(n, center=(0.6, 0.3, 0.5), radius=0.2))
    phi = O.liquid_sdf_postprocess(n3, dx, solid, O.liquid_sdf(n3, dx, pos, r))
    rng = np.random.default_rng(seed)
    vel = (rng.standard_normal(pos.shape) * 0.4).astype(np.float32)
    (u, v, w), valid = O.p2g(n3, dx, pos, vel, r)
    u, v, w = (O.extrapolate(f, vd, 12) for f, vd in zip((u, v, w), valid))
    return n3, O.pressure_system(n3, dx, 1.0 / 120.0, (u, v, w), O.face_weights(n3, solid), phi)


def _line_words(n3, rows):
    """lineWordsKernel restated: (base, bits) per 32-cell word of each line."""
    ni, nj, nk = n3
    nw = (ni + 31) // 32
    is_row = np.zeros(ni * nj * nk, bool)
    is_row[rows] = True
    row_of = np.concatenate([[0], np.cumsum(is_row)])[:-1]  # exclusive scan of the row mask
    lines = is_row.reshape(nk * nj, ni)
    bits = np.zeros((nk * nj, nw), np.uint64)
    base = np.zeros((nk * nj, nw), np.int64)
    for w in range(nw):
        seg = lines[:, 32 * w:32 * w + 32]
        bits[:, w] = (seg.astype(np.uint64) << np.arange(seg.shape[1], dtype=np.uint64)).sum(axis=1)
        base[:, w] = row_of.reshape(nk * nj, ni)[:, 32 * w]
    return nw, bits, base


def _popcount(v):
    v = v.astype(np.uint64)
    c = np.zeros(v.shape, np.int64)
    for b in range(32):
        c += ((v >> np.uint64(b)) & np.uint64(1)).astype(np.int64)
    return c


@pytest.mark.parametrize("n,seed", [(24, 1), (33, 2), (70, 3)])
def test_line_words_give_the_neighbour_rows(n, seed):
    n3, sysm = _system(n, seed)
    rows = sysm["rows"].astype(np.int64)
    nbr = sysm["nbr"]  # -x, +x, -y, +y, -z, +z row ids or -1
    m = len(rows)
    assert m > 0
    ni, nj, nk = n3
    nw, bits, base = _line_words(n3, rows)
    line, i = rows // ni, rows % ni
    bit = np.uint64(1) << (i % 32).astype(np.uint64)

    def word_row(dline):
        q = line + dline
        b, bs = bits[q, i // 32], base[q, i // 32]
        return np.where((b & bit) != 0, bs + _popcount(b & (bit - np.uint64(1))), -1)

    r = np.arange(m)
    prev = np.concatenate([[-10], rows[:-1]])
    nxt = np.concatenate([rows[1:], [-10]])
    assert np.array_equal(np.where(prev == rows - 1, r - 1, -1), nbr[:, 0])
    assert np.array_equal(np.where(nxt == rows + 1, r + 1, -1), nbr[:, 1])
    assert np.array_equal(word_row(-1), nbr[:, 2])
    assert np.array_equal(word_row(+1), nbr[:, 3])
    assert np.array_equal(word_row(-nj), nbr[:, 4])
    assert np.array_equal(word_row(+nj), nbr[:, 5])
