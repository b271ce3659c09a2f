"""MeshObject -- a triangle-mesh obstacle of the engine.

Same class, property names, validation and errors as the reference's
src/engine/ffengine/meshobject.py over the MeshObject C ABI
(c_bindings/meshobject_c.cpp).  This engine takes static and animated
meshes, normal or inversed, with friction.  The three meshes of
update_mesh_animated must have the same vertex and triangle counts (a
changing topology raises RuntimeError).
"""
from ctypes import c_double, c_float, c_int, c_void_p

from . import pybindings as pb
from .ffengine import ffengine as lib
from .fluidsimulation import Vector3, Vector3_t
from .trianglemesh import TriangleMesh_t


def _check_ge_zero(name, value):
    if value < 0:
        raise ValueError("%s must be greater than or equal to 0" % name)


class MeshObject():

    def __init__(self, i, j, k, dx):
        libfunc = lib.MeshObject_new
        pb.init_lib_func(libfunc, [c_int, c_int, c_int, c_double, c_void_p], c_void_p)
        self._obj = pb.execute_lib_func(libfunc, [i, j, k, dx])

    def __del__(self):
        try:
            libfunc = lib.MeshObject_destroy
            pb.init_lib_func(libfunc, [c_void_p], None)
            libfunc(self._obj)
        except Exception:
            pass

    def __call__(self):
        return self._obj

    def _call(self, name, args=(), argtypes=(), restype=None):
        libfunc = getattr(lib, name)
        pb.init_lib_func(libfunc, [c_void_p] + list(argtypes) + [c_void_p], restype)
        return pb.execute_lib_func(libfunc, [self()] + list(args))

    def update_mesh_static(self, mesh):
        self._mesh_struct = mesh.to_struct()
        self._call("MeshObject_update_mesh_static", [self._mesh_struct], [TriangleMesh_t])

    def update_mesh_animated(self, mesh_previous, mesh_current, mesh_next):
        s = [m.to_struct() for m in (mesh_previous, mesh_current, mesh_next)]
        self._call("MeshObject_update_mesh_animated", s, [TriangleMesh_t] * 3)

    @property
    def enable(self):
        return bool(self._call("MeshObject_is_enabled", restype=c_int))

    @enable.setter
    def enable(self, boolval):
        self._call("MeshObject_enable" if boolval else "MeshObject_disable")

    @property
    def inverse(self):
        return bool(self._call("MeshObject_is_inversed", restype=c_int))

    @inverse.setter
    def inverse(self, boolval):
        if bool(boolval) != self.inverse:
            self._call("MeshObject_inverse")

    @property
    def friction(self):
        return self._call("MeshObject_get_friction", restype=c_float)

    @friction.setter
    def friction(self, value):
        _check_ge_zero("friction", value)
        if value > 1.0:
            raise ValueError("friction must be less than or equal to 1.0")
        self._call("MeshObject_set_friction", [value], [c_float])

    @property
    def velocity_scale(self):
        return self._call("MeshObject_get_velocity_scale", restype=c_float)

    @velocity_scale.setter
    def velocity_scale(self, value):
        self._call("MeshObject_set_velocity_scale", [value], [c_float])

    @property
    def whitewater_influence(self):
        return self._call("MeshObject_get_whitewater_influence", restype=c_float)

    @whitewater_influence.setter
    def whitewater_influence(self, value):
        _check_ge_zero("whitewater_influence", value)
        self._call("MeshObject_set_whitewater_influence", [value], [c_float])

    @property
    def dust_emission_strength(self):
        return self._call("MeshObject_get_dust_emission_strength", restype=c_float)

    @dust_emission_strength.setter
    def dust_emission_strength(self, value):
        _check_ge_zero("dust_emission_strength", value)
        self._call("MeshObject_set_dust_emission_strength", [value], [c_float])

    @property
    def sheeting_strength(self):
        return self._call("MeshObject_get_sheeting_strength", restype=c_float)

    @sheeting_strength.setter
    def sheeting_strength(self, value):
        _check_ge_zero("sheeting_strength", value)
        self._call("MeshObject_set_sheeting_strength", [value], [c_float])

    @property
    def mesh_expansion(self):
        return self._call("MeshObject_get_mesh_expansion", restype=c_float)

    @mesh_expansion.setter
    def mesh_expansion(self, value):
        self._call("MeshObject_set_mesh_expansion", [value], [c_float])

    @property
    def enable_append_object_velocity(self):
        return bool(self._call("MeshObject_is_append_object_velocity_enabled", restype=c_int))

    @enable_append_object_velocity.setter
    def enable_append_object_velocity(self, boolval):
        self._call("MeshObject_enable_append_object_velocity" if boolval else
                   "MeshObject_disable_append_object_velocity")

    @property
    def object_velocity_influence(self):
        return self._call("MeshObject_get_object_velocity_influence", restype=c_float)

    @object_velocity_influence.setter
    def object_velocity_influence(self, value):
        self._call("MeshObject_set_object_velocity_influence", [value], [c_float])

    @property
    def priority(self):
        return self._call("MeshObject_get_priority", restype=c_int)

    @priority.setter
    def priority(self, n):
        self._call("MeshObject_set_priority", [int(n)], [c_int])

    @property
    def source_id(self):
        return self._call("MeshObject_get_source_id", restype=c_int)

    @source_id.setter
    def source_id(self, n):
        self._call("MeshObject_set_source_id", [int(n)], [c_int])

    @property
    def viscosity(self):
        return self._call("MeshObject_get_viscosity", restype=c_int)

    @viscosity.setter
    def viscosity(self, v):
        self._call("MeshObject_set_viscosity", [v], [c_float])

    @property
    def density(self):
        return self._call("MeshObject_get_density", restype=c_int)

    @density.setter
    def density(self, v):
        self._call("MeshObject_set_density", [v], [c_float])

    @property
    def lifetime(self):
        return self._call("MeshObject_get_lifetime", restype=c_int)

    @lifetime.setter
    def lifetime(self, v):
        self._call("MeshObject_set_lifetime", [v], [c_float])

    @property
    def lifetime_variance(self):
        return self._call("MeshObject_get_lifetime_variance", restype=c_int)

    @lifetime_variance.setter
    def lifetime_variance(self, v):
        self._call("MeshObject_set_lifetime_variance", [v], [c_float])

    def get_source_color(self):
        c = self._call("MeshObject_get_source_color", restype=Vector3_t)
        return Vector3(c.x, c.y, c.z)

    def set_source_color(self, r, g=None, b=None):
        if g is None and b is None:
            r, g, b = r
        self._call("MeshObject_set_source_color", [r, g, b], [c_double, c_double, c_double])
