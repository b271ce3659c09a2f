"""A differentiable run of FLIP substeps on the device (FFAMD_rollout_*).

Rollout(n3, dx, pos, ...) holds particles at the fixed positions pos (P, 3)
whose initial velocities are the trilinear sample of an initial MAC field.
forward(u0, v0, w0) runs the substeps in the engine's order
(fluidsimulation.cpp:10078-10121, without particle removal) and returns the
particle density on the cells after the last one, shaped (k, j, i).
backward(density_bar) returns the gradient of <density_bar, density> in the
initial face grids, recomputing substeps from particle checkpoints kept
every `checkpoint_every` substeps.  The liquid SDF and P2G flags of each
substep are those of the last forward with freeze=False and are held fixed
by the backward; the collision response is taken as the identity there.

The solid (the domain walls, or solid_nodes: a nodal solid SDF such as a
static MeshObject's) is frictionless: the velocity constraint takes friction
0, the reference's default (fluidsimulation.h:2227).  A solid whose objects
carry friction is not differentiated: solid_friction != 0 raises ValueError.
"""
from ctypes import byref, c_double, c_int, c_longlong, c_void_p

import numpy as np

from . import pybindings as pb
from .ffengine import ffengine as lib


def _fp(a):
    return a.ctypes.data_as(c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _call(name, argtypes, args, restype=None):
    f = getattr(lib, name)
    pb.init_lib_func(f, argtypes + [c_void_p], restype)
    ok = c_int()
    r = f(*args, byref(ok))
    pb.check_success(ok.value, name + " - ")
    return r


class Rollout:
    def __init__(self, n3, dx, pos, dt=1.0 / 120.0, substeps=1, checkpoint_every=1, picflip=0.05,
                 gravity=(0.0, -9.81, 0.0), cfl=5, tol=1e-9, max_iter=900, solid_nodes=None, solid_friction=0.0):
        if solid_friction != 0.0:
            raise ValueError("Rollout: the solid's friction must be 0 (the rollout's velocity constraint and its "
                             "adjoint are those of a frictionless solid)")
        self.n3 = tuple(int(a) for a in n3)
        self.dx = float(dx)
        self._pos = _f32(pos).reshape(-1, 3)
        self.np = len(self._pos)
        self.generation = 0  # counts forwards: a backward differentiates the last one
        ni, nj, nk = self.n3
        solid = _f32(solid_nodes) if solid_nodes is not None else None
        self._h = _call("FFAMD_rollout_new",
                        [c_int] * 3 + [c_double, c_int, c_void_p, c_void_p, c_double, c_int, c_int] + [c_double] * 6
                        + [c_int],
                        [ni, nj, nk, self.dx, self.np, _fp(self._pos), _fp(solid) if solid is not None else None,
                         float(dt), int(substeps), int(checkpoint_every), float(picflip), float(gravity[0]),
                         float(gravity[1]), float(gravity[2]), float(cfl), float(tol), int(max_iter)],
                        restype=c_void_p)

    def close(self):
        if getattr(self, "_h", None):
            f = lib.FFAMD_rollout_destroy
            pb.init_lib_func(f, [c_void_p], None)
            f(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shapes(self):
        ni, nj, nk = self.n3
        return [(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]

    def forward(self, u0, v0, w0, freeze=False):
        ni, nj, nk = self.n3
        grids = [_f32(a) for a in (u0, v0, w0)]
        for a, s in zip(grids, self._shapes()):
            if a.shape != s:
                raise ValueError("face grid shaped %s, expected %s" % (a.shape, s))
        rho = np.zeros((nk, nj, ni), np.float32)
        _call("FFAMD_rollout_forward", [c_void_p] * 4 + [c_int, c_void_p],
              [self._h, _fp(grids[0]), _fp(grids[1]), _fp(grids[2]), 1 if freeze else 0, _fp(rho)])
        self.generation += 1
        return rho

    def backward(self, density_bar):
        ni, nj, nk = self.n3
        rb = _f32(density_bar)
        if rb.shape != (nk, nj, ni):
            raise ValueError("density cotangent shaped %s, expected %s" % (rb.shape, (nk, nj, ni)))
        outs = [np.zeros(s, np.float32) for s in self._shapes()]
        _call("FFAMD_rollout_backward", [c_void_p] * 5, [self._h, _fp(rb)] + [_fp(o) for o in outs])
        return tuple(outs)

    def positions(self):
        out = np.zeros((self.np, 3), np.float32)
        _call("FFAMD_rollout_positions", [c_void_p, c_void_p], [self._h, _fp(out)])
        return out

    def stats(self):
        """(particles moved by the collision response, forward pressure
        iterations, adjoint pressure iterations) of the last forward/backward."""
        c, f, b = c_longlong(), c_longlong(), c_longlong()
        _call("FFAMD_rollout_stats", [c_void_p] * 4, [self._h, byref(c), byref(f), byref(b)])
        return c.value, f.value, b.value

    def set_reference_summation(self, threads):
        """FP64 dot products of every pressure solve in the reference's
        summation order for `threads` threads (blaswrapper.h:27-57); 0 restores
        the parallel tree."""
        _call("FFAMD_rollout_set_reference_summation", [c_void_p, c_int], [self._h, int(threads)])
