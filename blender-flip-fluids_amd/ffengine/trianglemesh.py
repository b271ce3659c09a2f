"""TriangleMesh -- a triangle mesh handed to the engine's C ABI.

Counterpart of src/engine/ffengine/trianglemesh.py (TriangleMesh_t and class
TriangleMesh): flat float32 vertex coordinates and int32 vertex indices, the
.bobj byte layout (vertex count, xyz floats, triangle count, index triples)
and the by-value TriangleMesh_t struct of MeshUtils (meshutils.h:34-39).
"""
import ctypes
import struct

import numpy as np


class TriangleMesh_t(ctypes.Structure):
    _fields_ = [("vertices", ctypes.c_void_p),
                ("triangles", ctypes.c_void_p),
                ("num_vertices", ctypes.c_int),
                ("num_triangles", ctypes.c_int)]


class TriangleMesh(object):
    def __init__(self, vertices=None, triangles=None):
        # own copies: translate/scale/apply_transform must not touch the caller's arrays
        self.vertices = np.zeros(0, np.float32) if vertices is None else \
            np.array(vertices, np.float32, copy=True).reshape(-1)
        self.triangles = np.zeros(0, np.int32) if triangles is None else \
            np.array(triangles, np.int32, copy=True).reshape(-1)

    @classmethod
    def from_bobj(cls, bobj_data):
        nv = struct.unpack_from("i", bobj_data, 0)[0]
        off = 4
        verts = np.frombuffer(bobj_data, np.float32, 3 * nv, off)
        off += 12 * nv
        nt = struct.unpack_from("i", bobj_data, off)[0]
        off += 4
        tris = np.frombuffer(bobj_data, np.int32, 3 * nt, off)
        return cls(verts.copy(), tris.copy())

    def to_bobj(self):
        return (struct.pack("i", len(self.vertices) // 3) + self.vertices.tobytes() +
                struct.pack("i", len(self.triangles) // 3) + self.triangles.tobytes())

    def to_struct(self):
        """The by-value struct; it keeps its arrays alive through _keep."""
        v = np.ascontiguousarray(self.vertices, np.float32)
        t = np.ascontiguousarray(self.triangles, np.int32)
        s = TriangleMesh_t(v.ctypes.data_as(ctypes.c_void_p), t.ctypes.data_as(ctypes.c_void_p),
                           len(v) // 3, len(t) // 3)
        s._keep = (v, t)
        return s

    def apply_transform(self, matrix_world):
        """Row-major 4x4 affine transform of every vertex, in double, stored as float."""
        m = np.asarray(matrix_world, np.float64).reshape(4, 4)
        v = self.vertices.reshape(-1, 3).astype(np.float64)
        out = v @ m[:3, :3].T + m[:3, 3]
        self.vertices = out.astype(np.float32).reshape(-1)

    def translate(self, tx, ty, tz):
        v = self.vertices.reshape(-1, 3)
        v += np.array([tx, ty, tz], np.float32)

    def scale(self, scale):
        self.vertices *= np.float32(scale)
