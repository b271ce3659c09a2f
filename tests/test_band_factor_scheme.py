"""The MIC(0) factor's band scheme (csrc/kernels_band.hip
micFactorBandKernel) restated on the CPU and checked against the oracle's
hyperplane-by-hyperplane factor (pcgsolver.h:68-205): workgroups own BT x BT
tiles of lines along i and recompute the K-1 halo lines in -y/-z; a halo line
dj + dk lines away computes band levels q <= K-1-dj-dk only; the previous
level's rows come from the workgroup's own lines, the hyperplane before the
band from the finished factor.  The restatement asserts that every exchanged
neighbour row is the row's true -x/-y/-z neighbour and that every row is
written by exactly one tile, and its invdiag must equal the oracle's bit for
bit (small tiles and deep bands exercise halos crossing several tiles).
"""
import math

import numpy as np
import pytest

from test_row_topology import _system


def band_factor(n3, S, K, BT):
    ni, nj, nk = n3
    rows = S["rows"].astype(np.int64); m = len(rows)
    cellRow = -np.ones(ni * nj * nk, np.int64); cellRow[rows] = np.arange(m)
    diag, ox, oy, oz, nbr = S["diag"], S["ox"], S["oy"], S["oz"], S["nbr"]
    inv = np.zeros(m)  # global invdiag
    nL = ni + nj + nk - 2
    E = BT + K - 1
    tilesJ, tilesK = (nj + BT - 1) // BT, (nk + BT - 1) // BT
    for l0 in range(0, nL, K):
        for tile in range(tilesJ * tilesK):
            J0, K0 = (tile % tilesJ) * BT, (tile // tilesJ) * BT
            val = np.zeros((2, E * E)); rowv = -np.ones((2, E * E), np.int64)
            prev = np.zeros(E * E); prevRow = -np.ones(E * E, np.int64)
            writes = {}
            for q in range(K):
                buf, pb = q & 1, (q & 1) ^ 1
                newval = np.zeros(E * E); newrow = -np.ones(E * E, np.int64)
                for tid in range(E * E):
                    ej, ek = tid % E, tid // E
                    j, k = J0 - (K - 1) + ej, K0 - (K - 1) + ek
                    lineIn = 0 <= j < nj and 0 <= k < nk
                    dj, dk = max(K - 1 - ej, 0), max(K - 1 - ek, 0)
                    owned = ej >= K - 1 and ek >= K - 1
                    i = l0 + q - j - k
                    ok = lineIn and q <= K - 1 - dj - dk and 0 <= i < ni
                    c = i + ni * (j + nj * k) if ok else 0
                    row = int(cellRow[c]) if ok else -1
                    v = 0.0
                    if row >= 0:
                        if q == 0:
                            rz = int(cellRow[c - ni * nj]) if k > 0 else -1
                            ry = int(cellRow[c - ni]) if j > 0 else -1
                            rx = int(cellRow[c - 1]) if i > 0 else -1
                            vz = inv[rz] if rz >= 0 else 0.0; vy = inv[ry] if ry >= 0 else 0.0; vx = inv[rx] if rx >= 0 else 0.0
                        else:
                            assert ek >= 1 and ej >= 1
                            rz, vz = int(rowv[pb][tid - E]), val[pb][tid - E]
                            ry, vy = int(rowv[pb][tid - 1]), val[pb][tid - 1]
                            rx, vx = int(prevRow[tid]), prev[tid]
                            # the exchanged rows are the true neighbours
                            assert rz == nbr[row, 4] and ry == nbr[row, 2] and rx == nbr[row, 0], (row, rz, ry, rx, nbr[row])
                        a = diag[row]
                        if a < 1e-9: a = 0.0
                        d = a
                        for kk, ik, js in ((rz, vz, 5), (ry, vy, 3), (rx, vx, 1)):
                            if kk < 0 or diag[kk] < 1e-9: continue
                            okx, oky, okz = ox[kk], oy[kk], oz[kk]
                            L = (okx if js == 1 else (oky if js == 3 else okz)) * ik
                            d -= L * L
                            miss = 0.0
                            if js != 1 and nbr[kk, 1] >= 0: miss += okx * ik
                            if js != 3 and nbr[kk, 3] >= 0: miss += oky * ik
                            if js != 5 and nbr[kk, 5] >= 0: miss += okz * ik
                            d -= 0.97 * L * miss
                        v = d if a == 0 else (1 / math.sqrt(a) if d < 0.25 * a else 1 / math.sqrt(d))
                        if owned:
                            assert row not in writes
                            writes[row] = v
                    newval[tid] = v; newrow[tid] = row
                val[buf] = newval; rowv[buf] = newrow
                prev = newval.copy(); prevRow = newrow.copy()
            for r, v in writes.items():
                inv[r] = v
    return inv



@pytest.mark.parametrize("n,seed,K,BT", [(12, 1, 5, 4), (14, 2, 13, 11)])
def test_band_factor_scheme_gives_the_oracle_factor(n, seed, K, BT):
    n3, S = _system(n, seed)
    got = band_factor(n3, S, K, BT)
    want = np.asarray(S["invdiag"], np.float64)
    assert len(want) > 0
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))
