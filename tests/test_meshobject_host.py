"""The MeshObject C ABI and the TriangleMesh binding on the host (no GPU).

MeshObject is host state only (meshobject_c.cpp): its properties round-trip
with the reference's clamping (friction clamped to [0, 1], influences and
strengths to >= 0), the Python layer raises ValueError where the
reference's does, and invalid meshes raise RuntimeError.  TriangleMesh keeps
the .bobj byte layout of src/engine/ffengine/trianglemesh.py.
"""
import ctypes as C
import struct

import numpy as np
import pytest

from ffengine import MeshObject, TriangleMesh, scenes
from ffengine.ffengine import ffengine as lib


def test_triangle_mesh_bobj_round_trip():
    v, t = scenes.sphere_mesh((0.5, 0.5, 0.5), 0.25, 1)
    m = TriangleMesh(v, t)
    data = m.to_bobj()
    assert struct.unpack_from("i", data, 0)[0] == len(v)
    assert struct.unpack_from("i", data, 4 + 12 * len(v))[0] == len(t)
    back = TriangleMesh.from_bobj(data)
    assert np.array_equal(back.vertices, v.reshape(-1)) and np.array_equal(back.triangles, t.reshape(-1))
    s = m.to_struct()
    assert s.num_vertices == len(v) and s.num_triangles == len(t)
    m.translate(1.0, 2.0, 3.0)
    assert np.allclose(m.vertices.reshape(-1, 3), v + np.array([1.0, 2.0, 3.0], np.float32))


def test_mesh_object_properties_round_trip():
    mo = MeshObject(16, 16, 16, 1.0 / 16)
    v, t = scenes.sphere_mesh((0.5, 0.5, 0.5), 0.25, 1)
    mo.update_mesh_static(TriangleMesh(v, t))
    assert mo.enable and not mo.inverse
    mo.enable = False
    assert not mo.enable
    mo.inverse = True
    assert mo.inverse
    mo.inverse = True  # already inversed: no toggle
    assert mo.inverse
    for name, value in (("friction", 0.5), ("velocity_scale", 2.0), ("whitewater_influence", 0.25),
                        ("dust_emission_strength", 3.0), ("sheeting_strength", 0.75), ("mesh_expansion", -0.1),
                        ("object_velocity_influence", 0.5)):
        setattr(mo, name, value)
        assert abs(getattr(mo, name) - value) < 1e-7, name
    mo.priority = 3
    mo.source_id = 7
    assert mo.priority == 3 and mo.source_id == 7
    mo.enable_append_object_velocity = True
    assert mo.enable_append_object_velocity
    mo.viscosity = 2.9  # the reference's C getter returns int
    assert mo.viscosity == 2
    mo.set_source_color(0.25, 0.5, 0.75)
    c = mo.get_source_color()
    assert (c.x, c.y, c.z) == (0.25, 0.5, 0.75)


def test_mesh_object_validation():
    mo = MeshObject(8, 8, 8, 0.125)
    with pytest.raises(ValueError):
        mo.friction = -0.1
    with pytest.raises(ValueError):
        mo.whitewater_influence = -1.0
    # the C setter clamps like MeshObject::setFriction (meshobject.cpp:340-345)
    err = C.c_int()
    lib.MeshObject_set_friction.argtypes = [C.c_void_p, C.c_float, C.c_void_p]
    lib.MeshObject_set_friction(mo(), 4.0, C.byref(err))
    assert err.value == 1 and mo.friction == 1.0
    v, t = scenes.sphere_mesh((0.5, 0.5, 0.5), 0.25, 0)
    with pytest.raises(RuntimeError):
        mo.update_mesh_static(TriangleMesh(v, np.array([[0, 1, len(v)]], np.int32)))
    mo.update_mesh_animated(TriangleMesh(v, t), TriangleMesh(v, t), TriangleMesh(v, t))
    with pytest.raises(RuntimeError):  # a topology that changes between frames
        mo.update_mesh_animated(TriangleMesh(v[:-1], t[:-1]), TriangleMesh(v, t), TriangleMesh(v, t))
    with pytest.raises(RuntimeError):
        MeshObject(0, 8, 8, 0.125)
