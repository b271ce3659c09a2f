"""The surface tension settings of FluidSimulation on the host (no GPU):
defaults, round trips and the reference's validation (fluidsimulation.cpp:
2604-2744: negative constants and non-positive condition numbers raise; the
Python layer raises ValueError first, src/engine/ffengine/fluidsimulation.py
2293-2305, 2376-2403)."""
import ctypes as C

import pytest

from ffengine import FluidSimulation
from ffengine.ffengine import ffengine as lib


def test_defaults_and_round_trips():
    sim = FluidSimulation(8, 8, 8, 0.125)
    assert sim.surface_tension == 0.0
    assert sim.surface_tension_condition_number == 10.0
    assert not sim.enable_smooth_surface_tension_kernel
    sim.surface_tension = 0.25
    sim.surface_tension_condition_number = 3.5
    sim.enable_smooth_surface_tension_kernel = True
    assert sim.surface_tension == 0.25
    assert sim.surface_tension_condition_number == 3.5
    assert sim.enable_smooth_surface_tension_kernel
    sim.enable_smooth_surface_tension_kernel = False
    assert not sim.enable_smooth_surface_tension_kernel


def test_validation():
    sim = FluidSimulation(8, 8, 8, 0.125)
    with pytest.raises(ValueError):
        sim.surface_tension = -0.5
    with pytest.raises(ValueError):
        sim.surface_tension_condition_number = 0.0
    err = C.c_int()
    f = lib.FluidSimulation_set_surface_tension
    f.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
    f(sim(), -1.0, C.byref(err))
    assert err.value == 0
    g = lib.FluidSimulation_set_surface_tension_condition_number
    g.argtypes = [C.c_void_p, C.c_double, C.c_void_p]
    g(sim(), -2.0, C.byref(err))
    assert err.value == 0
    msg = lib.CBindings_get_error_message
    msg.restype = C.c_char_p
    assert b"condition number must be greater than 0" in msg()
