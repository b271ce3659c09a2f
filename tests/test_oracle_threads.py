"""The oracle's threaded loops (liquid SDF and P2G per block, SpMV and vector
updates per row, G2P and advection per particle) give the same bits as its
serial run: each output is written by one thread in the serial order.  The
benchmark's CPU baseline runs the oracle threaded."""
import numpy as np

import oracle_lib as O
from ffengine import scenes


def _run(n, threads, frames):
    O.set_threads(threads)
    try:
        dx = 1.0 / n
        pos = scenes.dam_break_particles(n, seed=7)
        rng = np.random.default_rng(7)
        vel = (rng.standard_normal(pos.shape) * 0.3).astype(np.float32)
        sim = O.OracleSim((n, n, n), dx, tol=1e-8)
        sim.load(pos, vel)
        sim.initialize()
        out = []
        for _ in range(frames):
            sim.update(1.0 / 30.0)
            p, v = sim.particles()
            out.append((p.copy(), v.copy(), sim.pressure_iterations()))
        return out
    finally:
        O.set_threads(1)


def test_threaded_oracle_matches_serial_bits():
    serial = _run(28, 1, 2)
    threaded = _run(28, 5, 2)
    for (p1, v1, i1), (p5, v5, i5) in zip(serial, threaded):
        assert i1 == i5
        assert np.array_equal(p1.view(np.uint32), p5.view(np.uint32))
        assert np.array_equal(v1.view(np.uint32), v5.view(np.uint32))
