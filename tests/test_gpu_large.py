"""Grids whose population arrays hold 2^32 bytes or more.

The vector kernel addresses each population row through a buffer resource
whose 64-bit base is the row start and whose extent is the row, so no 32-bit
byte count or offset depends on the lattice size.  D3Q19 with FP64 storage
and FP32 compute at 1024 x 1024 x 512 puts exactly 2^32 bytes in every
population array (N = 2^29 cells, 163 GB for the two lattices): a resource
spanning a whole population would wrap its 32-bit num_records to 0 there, so
every load would return 0 and every store would be dropped without an
error.  Size-independent checks: mass conservation through the FP64
reduction, and the vector kernel against the one-cell-per-lane kernel (64-bit
element indexing) on the first and last x-planes of every population, read
straight from the device lattice.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from metalbm_amd import (Algorithm, Config, ForceType, ForcingSchemeType, KernelShape,
                         LatticeType, Precision, kernel_shape)

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _hip():
    for name in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        if os.path.sep not in name or os.path.exists(name):
            L = ctypes.CDLL(name)
            break
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipMemcpy.restype = ctypes.c_int
    return L


HIP_MEMCPY_D2H = 2


def test_vector_kernel_beyond_4gib_per_population():
    nx, ny, nz = 1024, 1024, 512
    base = dict(lattice=LatticeType.D3Q19, forcing_scheme=ForcingSchemeType.Guo,
                force=ForceType.Constant, nx=nx, ny=ny, nz=nz, tau=0.55,
                amplitude=(1e-6, 0, 0), storage=Precision.F64, compute=Precision.F32,
                shifted=True)
    N, Q, plane = nx * ny * nz, 19, ny * nz
    assert N * 8 == 2 ** 32
    assert kernel_shape(Config(**base)) == (KernelShape.VECTOR, 2)
    hip = _hip()
    out = []
    for kernel in (KernelShape.VECTOR, KernelShape.CELL):
        say(f"case large D3Q19 {nx}x{ny}x{nz} F64 storage / F32 compute, kernel {kernel.name}")
        with Algorithm(Config(kernel=int(kernel), **base)) as a:
            assert a.handle_config().kernel == int(kernel)
            a.init_noise(1.0, 1e-3, 42)
            a.iterate(1, store=True)
            m1, e1, _ = a.reduce()
            a.iterate(2, store=True)
            m3, e3, _ = a.reduce()
            ptr, elem = a.device_populations()
            assert elem == 8
            sl = np.empty((Q, 2, plane))
            for i in range(Q):
                for k, x in enumerate((0, nx - 1)):
                    off = (i * N + x * plane) * elem
                    rc = hip.hipMemcpy(sl[i, k].ctypes.data, ptr + off, plane * elem,
                                       HIP_MEMCPY_D2H)
                    assert rc == 0
        say(f"{kernel.name}: mass {m1:.12e} -> {m3:.12e}, energy {e1:.6e} -> {e3:.6e}")
        out.append((m1, m3, e1, e3, sl))
    (m1, m3, e1, e3, sv), (_, mc3, ec1, ec3, sc) = out
    # mass: rho0 = 1 per cell, conserved to FP32-compute round-off per cell
    assert abs(m1 - N) <= 1e-6 * N and abs(m3 - m1) <= 1e-9 * N
    assert abs(mc3 - m3) <= 1e-9 * N
    # energy of the noise field (at most 3 (1e-3)^2 / 6 per cell, decaying)
    assert 0.0 < e1 < 1e-6 and 0.0 < e3 < 1e-6
    assert abs(e1 - ec1) <= 1e-5 * ec1 and abs(e3 - ec3) <= 1e-5 * ec3
    # stored shifted populations g = f - w: the two kernel shapes agree to the
    # rounding of their FP32 arithmetic (|g| ~ 1e-4 here, FP32 unit 6e-8)
    scale = np.max(np.abs(sc))
    err = np.max(np.abs(sv - sc))
    say(f"vector vs cell on x = 0 and x = {nx - 1}: max |dg| {err:.3e} of max |g| {scale:.3e}")
    assert scale > 1e-5
    assert err <= 1e-5 * scale
