"""rand() seeding of the substep (FluidSimulation::_stepFluid calls
srand(frame + substep) first, fluidsimulation.cpp:10079; the obstacles' sign
tests then draw their jitter, meshutils.cpp:93-117).

- The oracle's restatement of the C library's generator (oracle_glibc_rand)
  returns exactly what the C library's rand() returns after srand(seed), for
  seeds that include 0 (taken as 1) and the largest int.
- The reference engine built from its sources reseeds at every substep: after
  one frame of one substep with a static sphere obstacle, whose rebuild takes
  6 draws (the domain box's sign test, then the sphere's), the next rand()
  is the 7th value after srand(0), whatever the caller seeded before.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from ffengine import scenes


@pytest.mark.parametrize("seed", [0, 1, 5, 7, 1234, 2 ** 31 - 1])
def test_glibc_restatement_equals_the_c_library(seed):
    n = 2000
    assert np.array_equal(O.glibc_rand_values(seed, n), O.libc_rand_values(seed, n))


@pytest.mark.skipif(not O.reference_available(), reason="reference build (oracle/_ref) not present")
def test_reference_reseeds_rand_at_every_substep():
    n = 16
    dx = 1.0 / n
    v, t = scenes.sphere_mesh((0.55, 0.3, 0.5), 0.2, 2)
    ref = O.ReferenceSim((n, n, n), dx, threads=1)
    ref.load(scenes.dam_break_particles(n, seed=0))
    ref.add_mesh_obstacle((n, n, n), dx, v, t)
    ref.initialize()
    O.libc_srand(4321)
    ref.update(1.0 / 240.0)
    nxt = C.CDLL(None).rand()
    assert nxt == int(O.glibc_rand_values(0, 7)[6])
    assert nxt != int(O.libc_rand_values(4321, 7)[6])
