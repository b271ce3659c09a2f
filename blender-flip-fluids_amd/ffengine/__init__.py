"""ffengine -- Python binding of the MI355X FLIP engine (libffengine.so).

Mirrors the reference engine's ctypes package (src/engine/ffengine/ of
rlguy/Blender-FLIP-Fluids 1.8.5): the same FluidSimulation class surface for
the per-substep path, the same error behaviour (RuntimeError "<func> - <msg>"),
loading a from-scratch HIP engine instead of the CPU one.
"""
from .fluidsimulation import FluidSimulation, FluidSimulationFrameStats_t  # noqa: F401
from . import kernels, scenes  # noqa: F401
from .meshobject import MeshObject  # noqa: F401
from .trianglemesh import TriangleMesh  # noqa: F401
