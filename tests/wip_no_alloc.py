"""FFAMD_extrapolate_noalloc (the layered extrapolation with its status
bytes in a static device array, no allocation) against FFAMD_extrapolate on
the same inputs, bit for bit: two calls in a row, the second on a smaller
array, so it also runs over the first call's leftover status bytes."""
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


def _noalloc(field_dev, valid_dev, shape, layers, stream):
    wk, wj, wi = shape
    f = lib.FFAMD_extrapolate_noalloc
    pb.init_lib_func(f, [ctypes.c_int] * 3 + [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.c_void_p], None)
    pb.execute_lib_func(f, [wi, wj, wk, field_dev, valid_dev, layers, stream])


def test_extrapolate_noalloc_matches_allocating_entry_point():
    torch = pytest.importorskip("torch")
    assert torch.cuda.is_available()
    stream = torch.cuda.current_stream().cuda_stream
    for shape, seed in (((36, 33, 41), 1), ((18, 21, 17), 2)):
        rng = np.random.default_rng(seed)
        field = rng.standard_normal(shape).astype(np.float32)
        valid = (rng.random(shape) < 0.08).astype(np.uint8)
        field[valid == 0] = 0.0
        want = K.extrapolate(field, valid, 12)
        dfield = torch.from_numpy(field.copy()).cuda()
        dvalid = torch.from_numpy(valid).cuda()
        torch.cuda.synchronize()
        say(f"extrapolate_noalloc shape={shape} seed={seed}")
        _noalloc(dfield.data_ptr(), dvalid.data_ptr(), shape, 12, stream)
        torch.cuda.synchronize()
        got = dfield.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), shape
    with pytest.raises(RuntimeError, match="at most"):
        _noalloc(dfield.data_ptr(), dvalid.data_ptr(), (1024, 1024, 65), 12, stream)
