"""FluidSimulation -- ctypes class over libffengine.so.

Same method and property names, argument meanings and errors as the
reference's src/engine/ffengine/fluidsimulation.py (class FluidSimulation,
fluidsimulation.py:54-3103) for the calls the per-substep path uses, plus
additive calls for the pressure precision/tolerance, an injected solid SDF
and grid read-back (numpy arrays, i fastest, shaped (k, j, i)).
"""
import collections
import ctypes
import numbers
from ctypes import (byref, c_char, c_char_p, c_double, c_float, c_int, c_uint, c_void_p)

import numpy as np

from . import pybindings as pb
from .ffengine import ffengine as lib

GridIndex = collections.namedtuple("GridIndex", "i j k")
Vector3 = collections.namedtuple("Vector3", "x y z")


class Vector3_t(ctypes.Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]


class FluidSimulationMarkerParticleData_t(ctypes.Structure):
    _fields_ = [("size", c_int), ("positions", c_char_p), ("velocities", c_char_p)]


class FluidSimulationMarkerParticleAffineData_t(ctypes.Structure):
    _fields_ = [("size", c_int), ("affinex", c_char_p), ("affiney", c_char_p), ("affinez", c_char_p)]


class FluidSimulationMeshStats_t(ctypes.Structure):
    _fields_ = [("enabled", c_int), ("vertices", c_int), ("triangles", c_int), ("bytes", c_uint)]


class FluidSimulationTimingStats_t(ctypes.Structure):
    _fields_ = [(n, c_double) for n in
                ("total", "mesh", "advection", "particles", "pressure", "diffuse", "viscosity", "objects")]


_MESH_STATS = [
    "surface", "preview", "surfaceblur", "surfacevelocity", "surfacevorticity", "surfacespeed",
    "surfaceage", "surfacelifetime", "surfacewhitewaterproximity", "surfacecolor", "surfacesourceid",
    "surfaceviscosity", "surfacedensity", "foam", "bubble", "spray", "dust", "foamblur", "bubbleblur",
    "sprayblur", "dustblur", "foamvelocity", "bubblevelocity", "sprayvelocity", "dustvelocity",
    "foamid", "bubbleid", "sprayid", "dustid", "foamlifetime", "bubblelifetime", "spraylifetime",
    "dustlifetime", "fluidparticles", "fluidparticlesid", "fluidparticlesuid", "fluidparticlesvelocity",
    "fluidparticlesspeed", "fluidparticlesvorticity", "fluidparticlescolor", "fluidparticlesage",
    "fluidparticleslifetime", "fluidparticlesviscosity", "fluidparticlesdensity",
    "fluidparticlesdensityaverage", "fluidparticleswhitewaterproximity", "fluidparticlessourceid",
    "particles", "obstacle", "forcefield",
]


class FluidSimulationFrameStats_t(ctypes.Structure):
    """Layout of FluidSimulationFrameStats (fluidsimulation.h:80-151)."""
    _fields_ = ([("frame", c_int), ("substeps", c_int), ("delta_time", c_double),
                 ("fluid_particles", c_int), ("diffuse_particles", c_int), ("performance_score", c_int),
                 ("pressure_solver_enabled", c_int), ("pressure_solver_success", c_int),
                 ("pressure_solver_error", c_double), ("pressure_solver_iterations", c_int),
                 ("pressure_solver_max_iterations", c_int), ("viscosity_solver_enabled", c_int),
                 ("viscosity_solver_success", c_int), ("viscosity_solver_error", c_double),
                 ("viscosity_solver_iterations", c_int), ("viscosity_solver_max_iterations", c_int)]
                + [(n, FluidSimulationMeshStats_t) for n in _MESH_STATS]
                + [("timing", FluidSimulationTimingStats_t)])


def _fp(a):
    return a.ctypes.data_as(c_void_p)


def _check_simulation_initialized(func):
    def wrapper(self, *args):
        if not self.is_initialized():
            raise RuntimeError("FluidSimulation must be initialized before calling this method")
        return func(self, *args)
    return wrapper


def _check_simulation_not_initialized(func):
    def wrapper(self, *args):
        if self.is_initialized():
            raise RuntimeError("This method must be called before FluidSimulation is initialized")
        return func(self, *args)
    return wrapper


def _int_getter(name):
    def get(self):
        f = getattr(lib, name)
        pb.init_lib_func(f, [c_void_p, c_void_p], c_int)
        return pb.execute_lib_func(f, [self()])
    return get


def _require_at_least(value, low):
    """method_decorators.check_ge_zero / check_ge (src/engine/ffengine/method_decorators.py)."""
    if isinstance(value, numbers.Real) and value < low:
        raise ValueError("Value must be greater than or equal to %s" % ("zero" if low == 0 else low))


def _setter(name, ctype):
    def set_(self, value):
        f = getattr(lib, name)
        pb.init_lib_func(f, [c_void_p, ctype, c_void_p], None)
        pb.execute_lib_func(f, [self(), value])
    return set_


def _call0(self, name, restype=None):
    f = getattr(lib, name)
    pb.init_lib_func(f, [c_void_p, c_void_p], restype)
    return pb.execute_lib_func(f, [self()])


class FluidSimulation(object):

    def __init__(self, isize=None, jsize=None, ksize=None, dx=None):
        is_dims = (isinstance(isize, int) and isinstance(jsize, int) and isinstance(ksize, int)
                   and isinstance(dx, numbers.Real))
        if not is_dims:
            raise TypeError("FluidSimulation must be initialized with types:\n"
                            "isize:\t%s\njsize:\t%s\nksize:\t%s\ndx:\t%s" % (int, int, int, float))
        if isize <= 0 or jsize <= 0 or ksize <= 0 or dx <= 0:
            raise ValueError("FluidSimulation dimensions and cell size must be greater than 0")
        f = lib.FluidSimulation_new_from_dimensions
        pb.init_lib_func(f, [c_int, c_int, c_int, c_double, c_void_p], c_void_p)
        self._obj = pb.execute_lib_func(f, [isize, jsize, ksize, dx])
        self._dims = (isize, jsize, ksize)

    def __del__(self):
        try:
            f = lib.FluidSimulation_destroy
            pb.init_lib_func(f, [c_void_p], None)
            f(self._obj)
        except Exception:
            pass
        self._obj = None

    def __call__(self):
        return self._obj

    # ---- lifecycle ----
    def initialize(self):
        if self.is_initialized():
            return
        _call0(self, "FluidSimulation_initialize")

    def is_initialized(self):
        return bool(_call0(self, "FluidSimulation_is_initialized", c_int))

    @_check_simulation_initialized
    def update(self, dt):
        f = lib.FluidSimulation_update
        pb.init_lib_func(f, [c_void_p, c_double, c_void_p], None)
        pb.execute_lib_func(f, [self(), dt])

    def get_current_frame(self):
        return _call0(self, "FluidSimulation_get_current_frame", c_int)

    def set_current_frame(self, frameno):
        _setter("FluidSimulation_set_current_frame", c_int)(self, frameno)

    def is_current_frame_finished(self):
        return bool(_call0(self, "FluidSimulation_is_current_frame_finished", c_int))

    def get_cell_size(self):
        return _call0(self, "FluidSimulation_get_cell_size", c_double)

    def get_grid_dimensions(self):
        f = lib.FluidSimulation_get_grid_dimensions
        pb.init_lib_func(f, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p], None)
        i, j, k, ok = c_int(), c_int(), c_int(), c_int()
        f(self(), byref(i), byref(j), byref(k), byref(ok))
        pb.check_success(ok.value, f.__name__ + " - ")
        return GridIndex(i.value, j.value, k.value)

    # ---- settings ----
    density = property(lambda s: _call0(s, "FluidSimulation_get_density", c_double),
                       _setter("FluidSimulation_set_density", c_double))
    pressure_solver_max_iterations = property(
        _int_getter("FluidSimulation_get_pressure_solver_max_iterations"),
        _setter("FluidSimulation_set_pressure_solver_max_iterations", c_int))
    max_thread_count = property(_int_getter("FluidSimulation_get_max_thread_count"),
                                _setter("FluidSimulation_set_max_thread_count", c_int))
    CFL_condition_number = property(_int_getter("FluidSimulation_get_CFL_condition_number"),
                                    _setter("FluidSimulation_set_CFL_condition_number", c_int))
    min_time_steps_per_frame = property(_int_getter("FluidSimulation_get_min_time_steps_per_frame"),
                                        _setter("FluidSimulation_set_min_time_steps_per_frame", c_int))
    max_time_steps_per_frame = property(_int_getter("FluidSimulation_get_max_time_steps_per_frame"),
                                        _setter("FluidSimulation_set_max_time_steps_per_frame", c_int))
    PICFLIP_ratio = property(lambda s: _call0(s, "FluidSimulation_get_PICFLIP_ratio", c_double),
                             _setter("FluidSimulation_set_PICFLIP_ratio", c_double))
    PICAPIC_ratio = property(lambda s: _call0(s, "FluidSimulation_get_PICAPIC_ratio", c_double),
                             _setter("FluidSimulation_set_PICAPIC_ratio", c_double))
    # constant viscosity and its solver (src/engine/ffengine/fluidsimulation.py
    # 335-346, 2267-2297): the first set enables the viscosity solve of every
    # substep; the reference's decorators raise ValueError below 0 (below 1
    # for the iteration limit)
    @property
    def viscosity(self):
        return _call0(self, "FluidSimulation_get_viscosity", c_double)

    @viscosity.setter
    def viscosity(self, value):
        _require_at_least(value, 0)
        _setter("FluidSimulation_set_viscosity", c_double)(self, value)

    # surface tension (src/engine/ffengine/fluidsimulation.py:2293-2305,
    # 2376-2403)
    @property
    def surface_tension(self):
        return _call0(self, "FluidSimulation_get_surface_tension", c_double)

    @surface_tension.setter
    def surface_tension(self, value):
        _require_at_least(value, 0)
        _setter("FluidSimulation_set_surface_tension", c_double)(self, value)

    @property
    def surface_tension_condition_number(self):
        return _call0(self, "FluidSimulation_get_surface_tension_condition_number", c_double)

    @surface_tension_condition_number.setter
    def surface_tension_condition_number(self, n):
        if isinstance(n, numbers.Real) and n <= 0.0:
            raise ValueError("Value must be greater than 0.0")
        _setter("FluidSimulation_set_surface_tension_condition_number", c_double)(self, n)

    @property
    def enable_smooth_surface_tension_kernel(self):
        return bool(_call0(self, "FluidSimulation_is_smooth_surface_tension_kernel_enabled", c_int))

    @enable_smooth_surface_tension_kernel.setter
    def enable_smooth_surface_tension_kernel(self, on):
        name = "FluidSimulation_%s_smooth_surface_tension_kernel" % ("enable" if on else "disable")
        f = getattr(lib, name)
        pb.init_lib_func(f, [c_void_p, c_void_p], None)
        pb.execute_lib_func(f, [self()])

    @property
    def viscosity_solver_error_tolerance(self):
        return _call0(self, "FluidSimulation_get_viscosity_solver_error_tolerance", c_double)

    @viscosity_solver_error_tolerance.setter
    def viscosity_solver_error_tolerance(self, value):
        _require_at_least(value, 0)
        _setter("FluidSimulation_set_viscosity_solver_error_tolerance", c_double)(self, value)

    @property
    def viscosity_solver_max_iterations(self):
        return _int_getter("FluidSimulation_get_viscosity_solver_max_iterations")(self)

    @viscosity_solver_max_iterations.setter
    def viscosity_solver_max_iterations(self, n):
        _require_at_least(n, 1)
        _setter("FluidSimulation_set_viscosity_solver_max_iterations", c_int)(self, int(n))

    @property
    def enable_surface_reconstruction(self):
        return bool(_call0(self, "FluidSimulation_is_surface_reconstruction_enabled", c_int))

    @enable_surface_reconstruction.setter
    def enable_surface_reconstruction(self, boolval):
        _call0(self, "FluidSimulation_enable_surface_reconstruction" if boolval
               else "FluidSimulation_disable_surface_reconstruction")

    @property
    def enable_extreme_velocity_removal(self):
        return bool(_call0(self, "FluidSimulation_is_extreme_velocity_removal_enabled", c_int))

    @enable_extreme_velocity_removal.setter
    def enable_extreme_velocity_removal(self, boolval):
        _call0(self, "FluidSimulation_enable_extreme_velocity_removal" if boolval
               else "FluidSimulation_disable_extreme_velocity_removal")

    def set_velocity_transfer_method_FLIP(self):
        _call0(self, "FluidSimulation_set_velocity_transfer_method_FLIP")

    def set_velocity_transfer_method_APIC(self):
        _call0(self, "FluidSimulation_set_velocity_transfer_method_APIC")

    def is_velocity_transfer_method_FLIP(self):
        return bool(_call0(self, "FluidSimulation_is_velocity_transfer_method_FLIP", c_int))

    def is_velocity_transfer_method_APIC(self):
        return bool(_call0(self, "FluidSimulation_is_velocity_transfer_method_APIC", c_int))

    def add_body_force(self, fx, fy, fz):
        f = lib.FluidSimulation_add_body_force
        pb.init_lib_func(f, [c_void_p, c_double, c_double, c_double, c_void_p], None)
        pb.execute_lib_func(f, [self(), fx, fy, fz])

    def get_constant_body_force(self):
        f = lib.FluidSimulation_get_constant_body_force
        pb.init_lib_func(f, [c_void_p, c_void_p], Vector3_t)
        ok = c_int()
        v = f(self(), byref(ok))
        pb.check_success(ok.value, f.__name__ + " - ")
        return Vector3(v.x, v.y, v.z)

    def reset_body_force(self):
        _call0(self, "FluidSimulation_reset_body_force")

    # ---- additive settings ----
    pressure_solver_tolerance = property(
        lambda s: _call0(s, "FluidSimulation_get_pressure_solver_tolerance", c_double),
        _setter("FluidSimulation_set_pressure_solver_tolerance", c_double))

    def set_pressure_solver_precision(self, mode, inner_tolerance=None):
        """mode 'fp64' (the reference's PCG), 'mixed' (FP32 PCG restarted inside
        FP64 iterative refinement; inner_tolerance = relative tolerance of each
        inner solve, default 1e-4) or 'mixed-reliable' (FP32 PCG with FP64
        reliable updates; inner_tolerance = FP32 residual drop between true
        residual updates, default 0.1) or 'fp32-preconditioner' (FP64 PCG
        whose MIC(0) application runs in FP32)."""
        m = {"fp64": 0, "mixed": 1, "mixed-reliable": 2, "fp32-preconditioner": 3}[mode] \
            if isinstance(mode, str) else int(mode)
        if inner_tolerance is None:
            inner_tolerance = 0.1 if m == 2 else 1e-4
        f = lib.FluidSimulation_set_pressure_solver_precision
        pb.init_lib_func(f, [c_void_p, c_int, c_double, c_void_p], None)
        pb.execute_lib_func(f, [self(), m, inner_tolerance])

    def set_pressure_solver_reference_summation(self, enabled):
        """FP64 solve: dot products in the reference's summation order (its
        BLAS::dot split by the max thread count) instead of the parallel tree."""
        f = lib.FluidSimulation_set_pressure_solver_reference_summation
        pb.init_lib_func(f, [c_void_p, c_int, c_void_p], None)
        pb.execute_lib_func(f, [self(), 1 if enabled else 0])

    def set_pressure_solver_warm_start(self, enabled):
        """FP64 PCG modes: start each solve from the previous substep's
        pressure instead of 0 (same stopping threshold)."""
        f = lib.FluidSimulation_set_pressure_solver_warm_start
        pb.init_lib_func(f, [c_void_p, c_int, c_void_p], None)
        pb.execute_lib_func(f, [self(), 1 if enabled else 0])

    def set_pressure_solver_preconditioner_schedule(self, schedule):
        """Schedule of the MIC(0) application: 2 = bands of hyperplanes,
        4 = the line pipeline (one launch per triangular sweep), 5 = the bands
        recorded once per solve as a HIP graph and replayed; same bits."""
        f = lib.FluidSimulation_set_pressure_solver_preconditioner_schedule
        pb.init_lib_func(f, [c_void_p, c_int, c_void_p], None)
        pb.execute_lib_func(f, [self(), int(schedule)])

    def get_pressure_solver_preconditioner_schedule(self):
        return int(_call0(self, "FluidSimulation_get_pressure_solver_preconditioner_schedule", c_int))

    def get_pressure_solver_warm_start(self):
        return bool(_call0(self, "FluidSimulation_get_pressure_solver_warm_start", c_int))

    def get_pressure_solver_reference_summation(self):
        return bool(_call0(self, "FluidSimulation_get_pressure_solver_reference_summation", c_int))

    def get_pressure_solver_precision(self):
        return ("fp64", "mixed", "mixed-reliable", "fp32-preconditioner")[
            _call0(self, "FluidSimulation_get_pressure_solver_precision", c_int)]

    @property
    def boundary_friction(self):
        return _call0(self, "FluidSimulation_get_boundary_friction", c_double)

    @boundary_friction.setter
    def boundary_friction(self, f):
        libfunc = lib.FluidSimulation_set_boundary_friction
        pb.init_lib_func(libfunc, [c_void_p, c_double, c_void_p], None)
        pb.execute_lib_func(libfunc, [self(), f])

    def add_mesh_obstacle(self, mesh_object):
        libfunc = lib.FluidSimulation_add_mesh_obstacle
        pb.init_lib_func(libfunc, [c_void_p, c_void_p, c_void_p], None)
        pb.execute_lib_func(libfunc, [self(), mesh_object()])

    def remove_mesh_obstacle(self, mesh_object):
        libfunc = lib.FluidSimulation_remove_mesh_obstacle
        pb.init_lib_func(libfunc, [c_void_p, c_void_p, c_void_p], None)
        pb.execute_lib_func(libfunc, [self(), mesh_object()])

    def remove_mesh_obstacles(self):
        _call0(self, "FluidSimulation_remove_mesh_obstacles")

    def get_last_solid_rand_draws(self):
        """rand() values the last solid SDF rebuild drew (sign-test jitter)."""
        return _call0(self, "FFAMD_get_last_solid_rand_draws", c_int)

    @_check_simulation_not_initialized
    def set_solid_phi(self, phi_nodes):
        i, j, k = self._dims
        a = np.ascontiguousarray(phi_nodes, dtype=np.float32).reshape(-1)
        if a.size != (i + 1) * (j + 1) * (k + 1):
            raise ValueError("solid phi must hold (isize+1)*(jsize+1)*(ksize+1) values")
        f = lib.FluidSimulation_set_solid_phi
        pb.init_lib_func(f, [c_void_p, c_void_p, c_void_p], None)
        pb.execute_lib_func(f, [self(), _fp(a)])

    # ---- particle I/O ----
    def load_marker_particle_data(self, num_particles, position_data, velocity_data):
        c_pos = (c_char * len(position_data)).from_buffer_copy(position_data)
        c_vel = (c_char * len(velocity_data)).from_buffer_copy(velocity_data)
        pdata = FluidSimulationMarkerParticleData_t()
        pdata.size = c_int(num_particles)
        pdata.positions = ctypes.cast(c_pos, c_char_p)
        pdata.velocities = ctypes.cast(c_vel, c_char_p)
        f = lib.FluidSimulation_load_marker_particle_data
        pb.init_lib_func(f, [c_void_p, FluidSimulationMarkerParticleData_t, c_void_p], None)
        pb.execute_lib_func(f, [self(), pdata])

    def get_num_marker_particles(self):
        return _call0(self, "FluidSimulation_get_num_marker_particles", c_int)

    def _get_output_data(self, size_name, data_name):
        size = _call0(self, size_name, c_uint)
        c_data = (c_char * size)()
        f = getattr(lib, data_name)
        pb.init_lib_func(f, [c_void_p, c_void_p, c_void_p], None)
        pb.execute_lib_func(f, [self(), c_data])
        return bytes(c_data)

    def get_marker_particle_position_data(self):
        return self._get_output_data("FluidSimulation_get_marker_particle_position_data_size",
                                     "FluidSimulation_get_marker_particle_position_data")

    def get_marker_particle_velocity_data(self):
        return self._get_output_data("FluidSimulation_get_marker_particle_velocity_data_size",
                                     "FluidSimulation_get_marker_particle_velocity_data")

    def _get_output_data_range(self, name, start_idx, end_idx, size_of_element):
        n = max(end_idx - start_idx, 0)
        c_data = (c_char * (n * size_of_element))()
        f = getattr(lib, name)
        pb.init_lib_func(f, [c_void_p, c_int, c_int, c_void_p, c_void_p], None)
        pb.execute_lib_func(f, [self(), start_idx, end_idx, c_data])
        return bytes(c_data)

    def get_marker_particle_position_data_range(self, start_idx, end_idx):
        return self._get_output_data_range("FluidSimulation_get_marker_particle_position_data_range",
                                           start_idx, end_idx, 12)

    def get_marker_particle_velocity_data_range(self, start_idx, end_idx):
        return self._get_output_data_range("FluidSimulation_get_marker_particle_velocity_data_range",
                                           start_idx, end_idx, 12)

    def get_marker_particle_affinex_data_range(self, start_idx, end_idx):
        return self._get_output_data_range("FluidSimulation_get_marker_particle_affinex_data_range",
                                           start_idx, end_idx, 12)

    def get_marker_particle_affiney_data_range(self, start_idx, end_idx):
        return self._get_output_data_range("FluidSimulation_get_marker_particle_affiney_data_range",
                                           start_idx, end_idx, 12)

    def get_marker_particle_affinez_data_range(self, start_idx, end_idx):
        return self._get_output_data_range("FluidSimulation_get_marker_particle_affinez_data_range",
                                           start_idx, end_idx, 12)

    def load_marker_particle_affine_data(self, num_particles, affinex_data, affiney_data, affinez_data):
        bufs = [(c_char * len(d)).from_buffer_copy(d) for d in (affinex_data, affiney_data, affinez_data)]
        adata = FluidSimulationMarkerParticleAffineData_t()
        adata.size = c_int(num_particles)
        adata.affinex = ctypes.cast(bufs[0], c_char_p)
        adata.affiney = ctypes.cast(bufs[1], c_char_p)
        adata.affinez = ctypes.cast(bufs[2], c_char_p)
        f = lib.FluidSimulation_load_marker_particle_affine_data
        pb.init_lib_func(f, [c_void_p, FluidSimulationMarkerParticleAffineData_t, c_void_p], None)
        pb.execute_lib_func(f, [self(), adata])

    def get_frame_stats_data(self):
        f = lib.FluidSimulation_get_frame_stats_data
        pb.init_lib_func(f, [c_void_p, c_void_p], FluidSimulationFrameStats_t)
        ok = c_int()
        stats = f(self(), byref(ok))
        pb.check_success(ok.value, f.__name__ + " - ")
        return stats

    # ---- additive read-back (numpy) ----
    def _grid(self, name, shape):
        a = np.zeros(shape, np.float32)
        f = getattr(lib, name)
        pb.init_lib_func(f, [c_void_p, c_void_p, c_void_p], None)
        pb.execute_lib_func(f, [self(), _fp(a)])
        return a

    def get_liquid_sdf(self):
        i, j, k = self._dims
        return self._grid("FluidSimulation_get_liquid_sdf_data", (k, j, i))

    def get_solid_sdf(self):
        i, j, k = self._dims
        return self._grid("FluidSimulation_get_solid_sdf_data", (k + 1, j + 1, i + 1))

    def get_pressure_grid(self):
        i, j, k = self._dims
        return self._grid("FluidSimulation_get_pressure_grid_data", (k, j, i))

    def get_velocity_field(self):
        i, j, k = self._dims
        u = np.zeros((k, j, i + 1), np.float32)
        v = np.zeros((k, j + 1, i), np.float32)
        w = np.zeros((k + 1, j, i), np.float32)
        f = lib.FluidSimulation_get_velocity_field_data
        pb.init_lib_func(f, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p], None)
        pb.execute_lib_func(f, [self(), _fp(u), _fp(v), _fp(w)])
        return u, v, w

    def get_solid_velocity_field(self):
        """The solid's face velocities of the last solid rebuild (moving
        obstacles; zero for a static solid)."""
        i, j, k = self._dims
        u = np.zeros((k, j, i + 1), np.float32)
        v = np.zeros((k, j + 1, i), np.float32)
        w = np.zeros((k + 1, j, i), np.float32)
        f = lib.FluidSimulation_get_solid_velocity_field
        pb.init_lib_func(f, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p], None)
        pb.execute_lib_func(f, [self(), _fp(u), _fp(v), _fp(w)])
        return u, v, w

    def get_last_pressure_solve_stats(self):
        f = lib.FluidSimulation_get_last_pressure_solve_stats
        pb.init_lib_func(f, [c_void_p] + [c_void_p] * 5, None)
        it, err, ok, m, s = c_int(), c_double(), c_int(), c_int(), c_int()
        f(self(), byref(it), byref(err), byref(ok), byref(m), byref(s))
        pb.check_success(s.value, f.__name__ + " - ")
        f = lib.FluidSimulation_get_last_pressure_solve_rhs_norm
        pb.init_lib_func(f, [c_void_p, c_void_p], c_double)
        rhs = pb.execute_lib_func(f, [self()])
        return {"iterations": it.value, "error": err.value, "success": bool(ok.value), "unknowns": m.value,
                "rhs_inf": rhs}

    def get_last_viscosity_solve_stats(self):
        f = lib.FluidSimulation_get_last_viscosity_solve_stats
        pb.init_lib_func(f, [c_void_p] + [c_void_p] * 5, None)
        it, err, ok, m, s = c_int(), c_double(), c_int(), c_int(), c_int()
        f(self(), byref(it), byref(err), byref(ok), byref(m), byref(s))
        pb.check_success(s.value, f.__name__ + " - ")
        return {"iterations": it.value, "error": err.value, "success": bool(ok.value), "unknowns": m.value}

    def get_last_substep_phase_ms(self, phase):
        f = lib.FluidSimulation_get_last_substep_phase_ms
        pb.init_lib_func(f, [c_void_p, c_int, c_void_p], c_double)
        return pb.execute_lib_func(f, [self(), phase])

    @_check_simulation_initialized
    def step(self, dt):
        """One substep of length dt (no frame split, no frame statistics)."""
        f = lib.FluidSimulation_step
        pb.init_lib_func(f, [c_void_p, c_double, c_void_p], None)
        pb.execute_lib_func(f, [self(), dt])
