"""numpy wrappers over the kernel-level entry points of libffengine.so.

Each runs one HIP kernel group of the FLIP substep on host arrays (copied to
and from the device) -- the parity tests compare them with the CPU oracle on
identical inputs.  Grids are returned shaped (k, j, i).
"""
from ctypes import byref, c_double, c_float, c_int, c_void_p

import numpy as np

from . import pybindings as pb
from .ffengine import ffengine as lib


def _fp(a):
    return a.ctypes.data_as(c_void_p)


def _run(name, argtypes, args, restype=None):
    f = getattr(lib, name)
    pb.init_lib_func(f, argtypes + [c_void_p], restype)
    ok = c_int()
    r = f(*args, byref(ok))
    pb.check_success(ok.value, name + " - ")
    return r


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def cell_sort(n3, dx, pos):
    ni, nj, nk = n3
    pos = _f32(pos).reshape(-1, 3)
    npart, ncell = len(pos), ni * nj * nk
    keys = np.zeros(npart, np.int32)
    counts = np.zeros(ncell, np.int32)
    offsets = np.zeros(ncell + 1, np.int32)
    perm = np.zeros(npart, np.int32)
    _run("FFAMD_cell_sort", [c_int] * 3 + [c_double, c_int] + [c_void_p] * 5,
         [ni, nj, nk, dx, npart, _fp(pos), _fp(keys), _fp(counts), _fp(offsets), _fp(perm)])
    return keys, counts, offsets, perm


def liquid_sdf(n3, dx, pos, radius):
    ni, nj, nk = n3
    pos = _f32(pos).reshape(-1, 3)
    phi = np.zeros((nk, nj, ni), np.float32)
    _run("FFAMD_liquid_sdf", [c_int] * 3 + [c_double, c_int, c_void_p, c_double, c_void_p],
         [ni, nj, nk, dx, len(pos), _fp(pos), radius, _fp(phi)])
    return phi


def p2g(n3, dx, pos, vel, radius):
    ni, nj, nk = n3
    pos = _f32(pos).reshape(-1, 3)
    vel = _f32(vel).reshape(-1, 3)
    u = np.zeros((nk, nj, ni + 1), np.float32)
    v = np.zeros((nk, nj + 1, ni), np.float32)
    w = np.zeros((nk + 1, nj, ni), np.float32)
    vu, vv, vw = (np.zeros(a.shape, np.uint8) for a in (u, v, w))
    _run("FFAMD_p2g", [c_int] * 3 + [c_double, c_int, c_void_p, c_void_p, c_double] + [c_void_p] * 6,
         [ni, nj, nk, dx, len(pos), _fp(pos), _fp(vel), radius, _fp(u), _fp(v), _fp(w), _fp(vu), _fp(vv), _fp(vw)])
    return (u, v, w), (vu, vv, vw)


def extrapolate(field, valid, layers):
    a = np.array(field, np.float32, copy=True, order="C")
    vb = np.ascontiguousarray(valid, dtype=np.uint8)
    wk, wj, wi = a.shape
    _run("FFAMD_extrapolate", [c_int] * 3 + [c_void_p, c_void_p, c_int], [wi, wj, wk, _fp(a), _fp(vb), layers])
    return a


def extrapolate_vjp(bar, valid, layers):
    """Adjoint of extrapolate for the same flags and layer count."""
    a = np.array(bar, np.float32, copy=True, order="C")
    vb = np.ascontiguousarray(valid, dtype=np.uint8)
    wk, wj, wi = a.shape
    _run("FFAMD_extrapolate_vjp", [c_int] * 3 + [c_void_p, c_void_p, c_int], [wi, wj, wk, _fp(a), _fp(vb), layers])
    return a


def solid_box_phi(n3, dx):
    ni, nj, nk = n3
    phi = np.zeros((nk + 1, nj + 1, ni + 1), np.float32)
    _run("FFAMD_solid_box_phi", [c_int] * 3 + [c_double, c_void_p], [ni, nj, nk, dx, _fp(phi)])
    return phi


def mesh_levelset_fast(cells3, dx, offset, vertices, triangles):
    """One mesh's fast signed distance level set on cells3 cells at node
    offset `offset` (exact band 3, crossing-count signs).  Returns (phi
    (k, j, i) nodal, rand() draws)."""
    isz, jsz, ksz = cells3
    v = _f32(vertices).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.int32).reshape(-1, 3)
    phi = np.zeros((ksz + 1, jsz + 1, isz + 1), np.float32)
    draws = _run("FFAMD_mesh_levelset_fast", [c_int] * 3 + [c_double] + [c_int] * 4 + [c_void_p, c_int] + [c_void_p] * 2,
                 [isz, jsz, ksz, dx, int(offset[0]), int(offset[1]), int(offset[2]), len(v), _fp(v), len(t), _fp(t),
                  _fp(phi)], c_int)
    return phi, draws


def solid_phi_meshes(n3, dx, meshes):
    """The domain box SDF with static mesh obstacles (vertices, triangles[,
    inversed[, expansion]]) unioned in.  Returns (phi (k, j, i) nodal, rand()
    draws)."""
    ni, nj, nk = n3
    vs = [_f32(m[0]).reshape(-1, 3) for m in meshes]
    ts = [np.ascontiguousarray(m[1], np.int32).reshape(-1, 3) for m in meshes]
    nv = np.array([len(v) for v in vs], np.int32)
    nt = np.array([len(t) for t in ts], np.int32)
    inv = np.array([1 if (len(m) > 2 and m[2]) else 0 for m in meshes], np.int32)
    exp = np.array([m[3] if len(m) > 3 else 0.0 for m in meshes], np.float32)
    verts = np.ascontiguousarray(np.concatenate(vs) if vs else np.zeros((0, 3), np.float32))
    tris = np.ascontiguousarray(np.concatenate(ts) if ts else np.zeros((0, 3), np.int32))
    phi = np.zeros((nk + 1, nj + 1, ni + 1), np.float32)
    draws = _run("FFAMD_solid_phi_meshes", [c_int] * 3 + [c_double, c_int] + [c_void_p] * 7,
                 [ni, nj, nk, dx, len(meshes), _fp(nv), _fp(verts), _fp(nt), _fp(tris), _fp(inv), _fp(exp), _fp(phi)],
                 c_int)
    return phi, draws


def solid_update_moving(n3, dx, objects, frame_dt, progress):
    """The solid update with moving obstacles.  objects: dicts with prev, cur,
    next (vertices), tris, and optional animated (default True), inversed,
    expansion, velocity_scale.  Returns (phi (k, j, i) nodal, (u, v, w) solid
    face velocities, rand() draws)."""
    ni, nj, nk = n3
    cat = lambda key: np.ascontiguousarray(np.concatenate([_f32(o[key]).reshape(-1, 3) for o in objects]),
                                           np.float32)
    prev, cur, nxt = cat("prev"), cat("cur"), cat("next")
    tris = np.ascontiguousarray(np.concatenate([np.asarray(o["tris"], np.int32).reshape(-1, 3) for o in objects]),
                                np.int32)
    nv = np.array([len(_f32(o["cur"]).reshape(-1, 3)) for o in objects], np.int32)
    nt = np.array([len(np.asarray(o["tris"]).reshape(-1, 3)) for o in objects], np.int32)
    flags = np.array([(1 if o.get("animated", True) else 0) | (2 if o.get("inversed") else 0) for o in objects],
                     np.int32)
    exp = np.array([o.get("expansion", 0.0) for o in objects], np.float32)
    vsc = np.array([o.get("velocity_scale", 1.0) for o in objects], np.float32)
    phi = np.zeros((nk + 1, nj + 1, ni + 1), np.float32)
    u = np.zeros((nk, nj, ni + 1), np.float32)
    v = np.zeros((nk, nj + 1, ni), np.float32)
    w = np.zeros((nk + 1, nj, ni), np.float32)
    draws = _run("FFAMD_solid_update_moving", [c_int] * 3 + [c_double, c_int] + [c_void_p] * 9 +
                 [c_double, c_float] + [c_void_p] * 4,
                 [ni, nj, nk, dx, len(objects), _fp(nv), _fp(nt), _fp(prev), _fp(cur), _fp(nxt), _fp(tris),
                  _fp(flags), _fp(exp), _fp(vsc), frame_dt, progress, _fp(phi), _fp(u), _fp(v), _fp(w)], c_int)
    return phi, (u, v, w), draws


def _visc_sizes(n3):
    ni, nj, nk = n3
    states = [(nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni)]
    volumes = [(nk, nj, ni), (nk, nj, ni + 1), (nk, nj + 1, ni), (nk + 1, nj, ni), (nk + 1, nj + 1, ni),
               (nk + 1, nj, ni + 1), (nk, nj + 1, ni + 1)]
    return states, volumes


VISC_VOLUMES = ("center", "U", "V", "W", "edgeU", "edgeV", "edgeW")


def _split(flat, shapes):
    out, q = [], 0
    for sh in shapes:
        n = int(np.prod(sh))
        out.append(flat[q:q + n].reshape(sh))
        q += n
    return out


def viscosity_system(n3, dx, dt, liquid_phi, solid_nodes, viscosity, uvw):
    """The viscosity stage's system built on the GPU (FFAMD_viscosity_system):
    a dict with m, states (u, v, w faces), volumes (by name), table,
    rowstart, cols, vals and rhs."""
    ni, nj, nk = n3
    sh_s, sh_v = _visc_sizes(n3)
    ins = [_f32(liquid_phi), _f32(solid_nodes), _f32(viscosity)] + [_f32(a) for a in uvw]
    argt = [c_int] * 3 + [c_double] * 2 + [c_void_p] * 14
    nnz = c_int()
    m = _run("FFAMD_viscosity_system", argt, [ni, nj, nk, dx, dt] + [_fp(a) for a in ins] + [byref(nnz)] +
             [None] * 7, c_int)
    ns = sum(int(np.prod(x)) for x in sh_s)
    nv = sum(int(np.prod(x)) for x in sh_v)
    bufs = [np.zeros(ns, np.uint8), np.zeros(nv, np.float32), np.zeros(ns, np.int32), np.zeros(m + 1, np.int32),
            np.zeros(nnz.value, np.int32), np.zeros(nnz.value, np.float32), np.zeros(m, np.float32)]
    _run("FFAMD_viscosity_system", argt, [ni, nj, nk, dx, dt] + [_fp(a) for a in ins] + [byref(nnz)] +
         [_fp(b) for b in bufs], c_int)
    states, volumes, table, rowstart, cols, vals, rhs = bufs
    return {"m": m, "states": _split(states, sh_s), "volumes": dict(zip(VISC_VOLUMES, _split(volumes, sh_v))),
            "table": table, "rowstart": rowstart, "cols": cols, "vals": vals, "rhs": rhs}


def viscosity_solve(n3, dx, dt, liquid_phi, solid_nodes, viscosity, uvw, tol=1e-4, max_iter=900):
    """The viscosity stage on the GPU (FFAMD_viscosity_solve): returns
    ((u, v, w), iterations, error, success, m, solution on the rows)."""
    ni, nj, nk = n3
    u, v, w = (np.array(a, np.float32, copy=True) for a in uvw)
    it, err, ok = c_int(), c_float(), c_int()
    m = _run("FFAMD_viscosity_system", [c_int] * 3 + [c_double] * 2 + [c_void_p] * 14,
             [ni, nj, nk, dx, dt, _fp(_f32(liquid_phi)), _fp(_f32(solid_nodes)), _fp(_f32(viscosity)), _fp(u),
              _fp(v), _fp(w), byref(c_int())] + [None] * 7, c_int)
    x = np.zeros(max(m, 1), np.float32)
    m2 = _run("FFAMD_viscosity_solve", [c_int] * 3 + [c_double] * 2 + [c_void_p] * 3 + [c_double, c_int] +
              [c_void_p] * 7,
              [ni, nj, nk, dx, dt, _fp(_f32(liquid_phi)), _fp(_f32(solid_nodes)), _fp(_f32(viscosity)), tol,
               max_iter, _fp(u), _fp(v), _fp(w), byref(it), byref(err), byref(ok), _fp(x)], c_int)
    assert m2 == m
    return (u, v, w), it.value, err.value, bool(ok.value), m, x[:m]


def curvature_grid(n3, dx, liquid_phi):
    """(surface phi, curvature) of the liquid SDF on the GPU (FFAMD_curvature_grid)."""
    ni, nj, nk = n3
    surf = np.zeros((nk, nj, ni), np.float32)
    kg = np.zeros((nk, nj, ni), np.float32)
    _run("FFAMD_curvature_grid", [c_int] * 3 + [c_double] + [c_void_p] * 3,
         [ni, nj, nk, dx, _fp(_f32(liquid_phi)), _fp(surf), _fp(kg)])
    return surf, kg


def tension_cluster_status(n3, liquid_phi):
    """The surface tension cluster status per cell (FFAMD_tension_cluster_status)."""
    ni, nj, nk = n3
    st = np.zeros((nk, nj, ni), np.int8)
    _run("FFAMD_tension_cluster_status", [c_int] * 3 + [c_void_p] * 2, [ni, nj, nk, _fp(_f32(liquid_phi)), _fp(st)])
    return st


def pressure_solve_tension(n3, dx, dt, uvw, weights, liquid_phi, k, tol=1e-9, max_iter=900, reference_threads=1):
    """FP64 PCG pressure solve with surface tension k and the gradient applied
    (FFAMD_pressure_solve_tension): (pressure, (u, v, w), flags, iterations,
    error, success)."""
    ni, nj, nk = n3
    u, v, w = (np.array(a, np.float32, copy=True) for a in uvw)
    flags = [np.zeros(a.shape, np.uint8) for a in (u, v, w)]
    p = np.zeros((nk, nj, ni), np.float32)
    it, err = c_int(), c_double()
    ok = _run("FFAMD_pressure_solve_tension", [c_int] * 3 + [c_double] * 2 + [c_void_p] * 7 +
              [c_double, c_double, c_int, c_int] + [c_void_p] * 6,
              [ni, nj, nk, dx, dt, _fp(u), _fp(v), _fp(w)] + [_fp(_f32(x)) for x in weights] +
              [_fp(_f32(liquid_phi)), k, tol, max_iter, reference_threads, _fp(p), byref(it), byref(err)] +
              [_fp(f) for f in flags], c_int)
    return p, (u, v, w), tuple(flags), it.value, err.value, bool(ok)


def cell_weights(n3, phi_nodes):
    """1 - the solid's volume fraction of each cell, (k, j, i)."""
    ni, nj, nk = n3
    p = _f32(phi_nodes)
    wc = np.zeros((nk, nj, ni), np.float32)
    _run("FFAMD_cell_weights", [c_int] * 3 + [c_void_p] * 2, [ni, nj, nk, _fp(p), _fp(wc)])
    return wc


def condition_solid_velocity(n3, liquid_phi, weights, solid_velocity):
    """The solid face velocities (copies) zeroed around isolated liquid pockets."""
    ni, nj, nk = n3
    p = _f32(liquid_phi)
    wu, wv, ww = (_f32(a) for a in weights)
    us, vs, ws = (np.array(a, np.float32, copy=True) for a in solid_velocity)
    _run("FFAMD_condition_solid_velocity", [c_int] * 3 + [c_void_p] * 7,
         [ni, nj, nk, _fp(p), _fp(wu), _fp(wv), _fp(ww), _fp(us), _fp(vs), _fp(ws)])
    return us, vs, ws


def face_weights(n3, phi_nodes):
    ni, nj, nk = n3
    p = _f32(phi_nodes)
    wu = np.zeros((nk, nj, ni + 1), np.float32)
    wv = np.zeros((nk, nj + 1, ni), np.float32)
    ww = np.zeros((nk + 1, nj, ni), np.float32)
    _run("FFAMD_face_weights", [c_int] * 3 + [c_void_p] * 4, [ni, nj, nk, _fp(p), _fp(wu), _fp(wv), _fp(ww)])
    return wu, wv, ww


def pressure_solve(n3, dx, dt, uvw, weights, liquid_phi, tol=1e-9, max_iter=900, mode=0):
    """Pressure solve + projection; returns (pressure, (u, v, w), valid, iterations, error, success)."""
    ni, nj, nk = n3
    u, v, w = (np.array(a, np.float32, copy=True, order="C") for a in uvw)
    wu, wv, ww = (_f32(a) for a in weights)
    phi = _f32(liquid_phi)
    p = np.zeros((nk, nj, ni), np.float32)
    vu, vv, vw = (np.zeros(a.shape, np.uint8) for a in (u, v, w))
    it, err = c_int(), c_double()
    ok = _run("FFAMD_pressure_solve",
              [c_int] * 3 + [c_double, c_double] + [c_void_p] * 7 + [c_double, c_int, c_int] + [c_void_p] * 6,
              [ni, nj, nk, dx, dt, _fp(u), _fp(v), _fp(w), _fp(wu), _fp(wv), _fp(ww), _fp(phi), tol, max_iter,
               mode, _fp(p), byref(it), byref(err), _fp(vu), _fp(vv), _fp(vw)], c_int)
    return p, (u, v, w), (vu, vv, vw), it.value, err.value, bool(ok)


def pressure_system(n3, dx, dt, uvw, weights, liquid_phi):
    ni, nj, nk = n3
    m_cap = max(ni - 2, 1) * max(nj - 2, 1) * max(nk - 2, 1)
    u, v, w = (_f32(a) for a in uvw)
    wu, wv, ww = (_f32(a) for a in weights)
    phi = _f32(liquid_phi)
    rows = np.zeros(m_cap, np.int32)
    arr = [np.zeros(m_cap, np.float64) for _ in range(5)]
    nbr = np.zeros(6 * m_cap, np.int32)
    inv = np.zeros(m_cap, np.float64)
    m = _run("FFAMD_pressure_system", [c_int] * 3 + [c_double, c_double] + [c_void_p] * 15,
             [ni, nj, nk, dx, dt, _fp(u), _fp(v), _fp(w), _fp(wu), _fp(wv), _fp(ww), _fp(phi), _fp(rows)]
             + [_fp(a) for a in arr] + [_fp(nbr), _fp(inv)], c_int)
    rhs, diag, ox, oy, oz = (a[:m] for a in arr)
    return {"rows": rows[:m], "rhs": rhs, "diag": diag, "ox": ox, "oy": oy, "oz": oz,
            "nbr": nbr[:6 * m].reshape(m, 6), "invdiag": inv[:m]}


def mic0_apply(n3, dx, dt, uvw, weights, liquid_phi, r, precision=0, schedule=0):
    """One MIC(0) application z = (L L^T)^-1 r on the pressure system of these
    inputs (row order as pressure_system); precision 0 = FP64, 1 = FP32;
    schedule 0 = one launch per hyperplane, 1 = tile wavefront (switched off),
    2 = bands of hyperplanes, 3 = bands with the tile check on the device,
    4 = the line pipeline (one launch per triangular sweep)."""
    ni, nj, nk = n3
    u, v, w = (_f32(a) for a in uvw)
    wu, wv, ww = (_f32(a) for a in weights)
    phi = _f32(liquid_phi)
    r = np.ascontiguousarray(r, np.float64)
    z = np.zeros_like(r)
    _run("FFAMD_mic0_apply", [c_int] * 3 + [c_double, c_double] + [c_void_p] * 7 + [c_int] * 3 + [c_void_p] * 2,
         [ni, nj, nk, dx, dt, _fp(u), _fp(v), _fp(w), _fp(wu), _fp(wv), _fp(ww), _fp(phi), precision, schedule,
          len(r), _fp(r), _fp(z)], c_int)
    return z


def sample_velocity(n3, dx, uvw, pos):
    ni, nj, nk = n3
    u, v, w = (_f32(a) for a in uvw)
    pos = _f32(pos).reshape(-1, 3)
    out = np.zeros_like(pos)
    _run("FFAMD_sample_velocity", [c_int] * 3 + [c_double] + [c_void_p] * 3 + [c_int, c_void_p, c_void_p],
         [ni, nj, nk, dx, _fp(u), _fp(v), _fp(w), len(pos), _fp(pos), _fp(out)])
    return out


def projection_vjp(n3, dx, dt, weights, liquid_phi, grads, tol=1e-12, max_iter=5000):
    """Adjoint of the pressure projection (pressure_solve's velocity update)
    for these weights and liquid SDF: returns (u, v, w) = P^T (grads)."""
    ni, nj, nk = n3
    wu, wv, ww = (_f32(a) for a in weights)
    phi = _f32(liquid_phi)
    gu, gv, gw = (_f32(a) for a in grads)
    ou, ov, ow = (np.zeros(a.shape, np.float32) for a in (gu, gv, gw))
    it = c_int()
    ok = _run("FFAMD_projection_vjp",
              [c_int] * 3 + [c_double, c_double] + [c_void_p] * 7 + [c_double, c_int] + [c_void_p] * 4,
              [ni, nj, nk, dx, dt, _fp(wu), _fp(wv), _fp(ww), _fp(phi), _fp(gu), _fp(gv), _fp(gw), tol, max_iter,
               _fp(ou), _fp(ov), _fp(ow), byref(it)], c_int)
    return (ou, ov, ow), it.value, bool(ok)


def p2g_vjp(n3, dx, pos, radius, grads, vel=None):
    """Adjoint of p2g: vbar (P, 3), the velocities' cotangent; with vel also
    xbar (P, 3), the positions' cotangent, returned as (vbar, xbar)."""
    ni, nj, nk = n3
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    gu, gv, gw = (_f32(a) for a in grads)
    vbar = np.zeros_like(pos)
    xbar = np.zeros_like(pos) if vel is not None else None
    v = np.ascontiguousarray(vel, np.float32).reshape(-1, 3) if vel is not None else None
    _run("FFAMD_p2g_vjp", [c_int] * 3 + [c_double, c_int, c_void_p, c_void_p, c_double] + [c_void_p] * 5,
         [ni, nj, nk, dx, len(pos), _fp(pos), _fp(v) if v is not None else None, radius, _fp(gu), _fp(gv), _fp(gw),
          _fp(vbar), _fp(xbar) if xbar is not None else None])
    return vbar if vel is None else (vbar, xbar)


def sample_velocity_vjp(n3, dx, pos, vbar):
    """Adjoint of sample_velocity in the face grids (positions fixed): (u, v, w)."""
    ni, nj, nk = n3
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    vb = np.ascontiguousarray(vbar, np.float32).reshape(-1, 3)
    u = np.zeros((nk, nj, ni + 1), np.float32)
    v = np.zeros((nk, nj + 1, ni), np.float32)
    w = np.zeros((nk + 1, nj, ni), np.float32)
    _run("FFAMD_sample_velocity_vjp", [c_int] * 3 + [c_double, c_int] + [c_void_p] * 5,
         [ni, nj, nk, dx, len(pos), _fp(pos), _fp(vb), _fp(u), _fp(v), _fp(w)])
    return u, v, w


def advect_rk3(n3, dx, uvw, pos, dt):
    """RK3 advection of positions through the MAC field, no collision."""
    ni, nj, nk = n3
    u, v, w = (_f32(a) for a in uvw)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    out = np.zeros_like(pos)
    _run("FFAMD_advect_rk3", [c_int] * 3 + [c_double] + [c_void_p] * 3 + [c_int, c_void_p, c_double, c_void_p],
         [ni, nj, nk, dx, _fp(u), _fp(v), _fp(w), len(pos), _fp(pos), dt, _fp(out)])
    return out


def advect_rk3_vjp(n3, dx, uvw, pos, dt, xbar):
    """Adjoint of advect_rk3: (pbar (P, 3), (u, v, w) cotangents)."""
    ni, nj, nk = n3
    u, v, w = (_f32(a) for a in uvw)
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    xb = np.ascontiguousarray(xbar, np.float32).reshape(-1, 3)
    pbar = np.zeros_like(pos)
    gu, gv, gw = (np.zeros(a.shape, np.float32) for a in (u, v, w))
    _run("FFAMD_advect_rk3_vjp",
         [c_int] * 3 + [c_double] + [c_void_p] * 3 + [c_int, c_void_p, c_double] + [c_void_p] * 5,
         [ni, nj, nk, dx, _fp(u), _fp(v), _fp(w), len(pos), _fp(pos), dt, _fp(xb), _fp(pbar), _fp(gu), _fp(gv),
          _fp(gw)])
    return pbar, (gu, gv, gw)


def sample_velocity_position_vjp(n3, dx, uvw, pos, vbar):
    """Position cotangent of sample_velocity: xbar (P, 3) = grad V(pos)^T vbar."""
    ni, nj, nk = n3
    u, v, w = (_f32(a) for a in uvw)
    pos = _f32(pos).reshape(-1, 3)
    vb = _f32(vbar).reshape(-1, 3)
    xbar = np.zeros_like(pos)
    _run("FFAMD_sample_velocity_position_vjp", [c_int] * 3 + [c_double] + [c_void_p] * 3 + [c_int] + [c_void_p] * 3,
         [ni, nj, nk, dx, _fp(u), _fp(v), _fp(w), len(pos), _fp(pos), _fp(vb), _fp(xbar)])
    return xbar


def p2g_apic(n3, dx, pos, vel, aff, radius):
    """APIC P2G: ((u, v, w), (valid_u, valid_v, valid_w)); aff (P, 9) holds the
    AFFINEX, AFFINEY and AFFINEZ vectors of each particle."""
    ni, nj, nk = n3
    pos = _f32(pos).reshape(-1, 3)
    vel = _f32(vel).reshape(-1, 3)
    aff = _f32(aff).reshape(-1, 9)
    u = np.zeros((nk, nj, ni + 1), np.float32)
    v = np.zeros((nk, nj + 1, ni), np.float32)
    w = np.zeros((nk + 1, nj, ni), np.float32)
    vu, vv, vw = (np.zeros(a.shape, np.uint8) for a in (u, v, w))
    _run("FFAMD_p2g_apic", [c_int] * 3 + [c_double, c_int] + [c_void_p] * 3 + [c_double] + [c_void_p] * 6,
         [ni, nj, nk, dx, len(pos), _fp(pos), _fp(vel), _fp(aff), radius, _fp(u), _fp(v), _fp(w), _fp(vu), _fp(vv),
          _fp(vw)])
    return (u, v, w), (vu, vv, vw)


def g2p_apic(n3, dx, uvw, pos):
    """APIC G2P: (velocities (P, 3), affine vectors (P, 9))."""
    ni, nj, nk = n3
    u, v, w = (_f32(a) for a in uvw)
    pos = _f32(pos).reshape(-1, 3)
    vel = np.zeros_like(pos)
    aff = np.zeros((len(pos), 9), np.float32)
    _run("FFAMD_g2p_apic", [c_int] * 3 + [c_double] + [c_void_p] * 3 + [c_int] + [c_void_p] * 3,
         [ni, nj, nk, dx, _fp(u), _fp(v), _fp(w), len(pos), _fp(pos), _fp(vel), _fp(aff)])
    return vel, aff
