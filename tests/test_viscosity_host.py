"""The viscosity settings of FluidSimulation on the host (no GPU): defaults,
round trips and the reference's validation (fluidsimulation.cpp:303-312,
2565-2601: negative viscosity and tolerance raise in the engine; the Python
layer's decorators raise ValueError first, src/engine/ffengine/
fluidsimulation.py:335-346, 2267-2297)."""
import ctypes as C

import pytest

from ffengine import FluidSimulation
from ffengine.ffengine import ffengine as lib


def test_defaults_and_round_trips():
    sim = FluidSimulation(8, 8, 8, 0.125)
    assert sim.viscosity == 0.0
    assert sim.viscosity_solver_error_tolerance == 1e-4
    assert sim.viscosity_solver_max_iterations == 900
    sim.viscosity = 0.75
    sim.viscosity_solver_error_tolerance = 1e-6
    sim.viscosity_solver_max_iterations = 50
    assert sim.viscosity == 0.75
    assert sim.viscosity_solver_error_tolerance == 1e-6
    assert sim.viscosity_solver_max_iterations == 50
    st = sim.get_frame_stats_data()
    assert st.viscosity_solver_enabled == 0  # no frame yet


def test_python_layer_validation():
    sim = FluidSimulation(8, 8, 8, 0.125)
    with pytest.raises(ValueError):
        sim.viscosity = -0.1
    with pytest.raises(ValueError):
        sim.viscosity_solver_error_tolerance = -1e-3
    with pytest.raises(ValueError):
        sim.viscosity_solver_max_iterations = 0
    assert sim.viscosity == 0.0


def test_engine_validation_through_the_c_abi():
    sim = FluidSimulation(8, 8, 8, 0.125)
    err = C.c_int()
    f = lib.FluidSimulation_set_viscosity
    f.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
    f(sim(), -2.0, C.byref(err))
    assert err.value == 0
    msg = lib.CBindings_get_error_message
    msg.restype = C.c_char_p
    assert b"viscosity must be greater than or equal to 0" in msg()
    g = lib.FluidSimulation_set_viscosity_solver_error_tolerance
    g.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
    g(sim(), -1.0, C.byref(err))
    assert err.value == 0
    f(sim(), 3.0, C.byref(err))
    assert err.value == 1 and sim.viscosity == 3.0
