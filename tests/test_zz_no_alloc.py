"""FFAMD_extrapolate_noalloc (the layered extrapolation with its status
bytes in a static device array, no allocation) against FFAMD_extrapolate on
the same inputs, bit for bit: two calls in a row, the second on a smaller
array, so it also runs over the first call's leftover status bytes.

The device buffers come from the HIP runtime the engine library is linked
against (hipMalloc through ctypes), on the null stream: this test allocates,
the entry point under test does not."""
import ctypes
import sys

import numpy as np
import pytest

from ffengine import kernels as K
from ffengine import pybindings as pb
from ffengine.ffengine import ffengine as lib

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libffengine.so loaded."""

    def __init__(self):
        lib.path  # the engine library is loaded: its HIP runtime is the process's
        self.L = ctypes.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed with HIP error %d" % (what, rc))

    def upload(self, a):
        p = ctypes.c_void_p()
        self.check(self.L.hipMalloc(ctypes.byref(p), a.nbytes), "hipMalloc")
        self.check(self.L.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1), "hipMemcpy H2D")
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        self.check(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self.check(self.L.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self, p):
        self.check(self.L.hipFree(p), "hipFree")


def _noalloc(field_dev, valid_dev, shape, layers):
    wk, wj, wi = shape
    f = lib.FFAMD_extrapolate_noalloc
    pb.init_lib_func(f, [ctypes.c_int] * 3 + [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.c_void_p], None)
    pb.execute_lib_func(f, [wi, wj, wk, field_dev, valid_dev, layers, None])


def test_extrapolate_noalloc_matches_allocating_entry_point():
    hip = _Hip()
    for shape, seed in (((36, 33, 41), 1), ((18, 21, 17), 2)):
        rng = np.random.default_rng(seed)
        field = rng.standard_normal(shape).astype(np.float32)
        valid = (rng.random(shape) < 0.08).astype(np.uint8)
        field[valid == 0] = 0.0
        want = K.extrapolate(field, valid, 12)
        dfield, dvalid = hip.upload(field), hip.upload(valid)
        say(f"extrapolate_noalloc shape={shape} seed={seed}")
        _noalloc(dfield, dvalid, shape, 12)
        got = hip.download(dfield, field)
        hip.free(dfield)
        hip.free(dvalid)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), shape
    with pytest.raises(RuntimeError, match="at most"):
        _noalloc(None, None, (1024, 1024, 65), 12)
