"""The line-pipeline MIC(0) sweeps (csrc/pencil_sweep.h) emulated on the CPU.

tests/emu/pencil_emu.cpp compiles the GPU sweep's own body with g++ under
AddressSanitizer and UndefinedBehaviorSanitizer and runs it thread by thread
(one OS thread per lane, a std::barrier for each workgroup barrier, the
workgroups in ticket order).  Each case applies the preconditioner twice in a
row (the second application reads exchange words the first one left behind)
and must equal, bit for bit:
  * FP64: the oracle's serial sweeps (pcgsolver.h:211-241);
  * FP32: a hyperplane-vectorised FP32 restatement of the same sweeps.
Every array access of the sweep's indexing is checked against the buffers'
extents by the sanitizer.  This runs before the kernel's GPU test.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))


def _inputs(n):
    # the GPU kernel tests' pressure scene (import here: that module is GPU-marked)
    from test_gpu_kernels import _pressure_inputs
    return _pressure_inputs(n)


def _mic_sweeps(sysm, n3, r, dtype):
    from test_gpu_kernels import _mic_sweeps as sweeps
    return sweeps(sysm, n3, r, dtype)


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    exe = str(tmp_path_factory.mktemp("emu") / "pencil_emu")
    subprocess.run(["g++", "-std=c++20", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-pthread", os.path.join(HERE, "emu", "pencil_emu.cpp"), "-o", exe],
                   check=True)
    return exe


def _records(sysm):
    """Forward {inv, L of the -x/-y/-z rows toward the row} and backward
    {inv, the row's own L toward +x/+y/+z} records, as mic0LowerEntries
    writes them."""
    null = sysm["diag"] < 1e-9
    inv = sysm["invdiag"]
    lx, ly, lz = (np.where(null, sysm[k], sysm[k] * inv) for k in ("ox", "oy", "oz"))
    nb = sysm["nbr"]
    def lower(l, slot):
        k = nb[:, slot]
        return np.where(k >= 0, l[np.maximum(k, 0)], 0.0)
    fwd = np.stack([inv, lower(lx, 0), lower(ly, 2), lower(lz, 4)], axis=1)
    bwd = np.stack([inv, lx, ly, lz], axis=1)
    return fwd, bwd


def _run(exe, tmp_path, n3, sysm, r, fp64):
    ni, nj, nk = n3
    N = ni * nj * nk
    m = len(sysm["rows"])
    cell_row = np.full(N, -1, np.int32)
    cell_row[sysm["rows"]] = np.arange(m, dtype=np.int32)
    row_of = np.zeros(N + 1, np.int32)
    row_of[1:] = np.cumsum(cell_row >= 0)
    fwd, bwd = _records(sysm)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        f.write(np.array([ni, nj, nk, m, 1 if fp64 else 0], np.int32).tobytes())
        f.write(cell_row.tobytes())
        f.write(row_of.tobytes())
        f.write(np.ascontiguousarray(r, np.float64).tobytes())
        f.write(np.ascontiguousarray(fwd, np.float64).tobytes())
        f.write(np.ascontiguousarray(bwd, np.float64).tobytes())
    subprocess.run([exe, str(fin), str(fout)], check=True, timeout=600)
    return np.fromfile(fout, np.float64)


@pytest.mark.parametrize("n", [24, 33, 70])
def test_pencil_sweeps_emulated_bitexact(emulator, tmp_path, n):
    n3, dx, dt, uvw, weights, phi = _inputs(n)
    sysm = O.pressure_system(n3, dx, dt, uvw, weights, phi)
    m = len(sysm["rows"])
    assert m > 0
    r = np.random.default_rng(n).standard_normal(m)
    z64 = _run(emulator, tmp_path, n3, sysm, r, fp64=True)
    assert np.array_equal(z64.view(np.uint64), O.apply_mic0(sysm, r).view(np.uint64))
    z32 = _run(emulator, tmp_path, n3, sysm, r, fp64=False).astype(np.float32)
    want = _mic_sweeps(sysm, n3, r, np.float32)
    assert np.array_equal(z32.view(np.uint32), want.view(np.uint32))
