"""The C ABI library loads (no GPU needed) and exports every entry point that
include/ffengine_amd.h declares."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ffengine_amd.h")
LIB = os.path.join(ROOT, "blender-flip-fluids_amd", "lib", "libffengine.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b((?:FluidSimulation|FFAMD|CBindings)_\w+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_path():
    names = declared_functions()
    for required in ("FluidSimulation_new_from_dimensions", "FluidSimulation_destroy", "FluidSimulation_initialize",
                     "FluidSimulation_update", "FluidSimulation_load_marker_particle_data",
                     "FluidSimulation_get_marker_particle_position_data_range", "CBindings_get_error_message"):
        assert required in names


def test_library_exports_every_declared_symbol():
    assert os.path.isfile(LIB), "build libffengine.so first"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
