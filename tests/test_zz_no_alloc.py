"""lbm_reduce_noalloc (no device allocation, static scratch) against
lbm_reduce, called through the C ABI: bit for bit on two calls in a row, the
second on a smaller grid, and a grid beyond the static scratch is rejected
with a clear error."""
import ctypes

import numpy as np
import pytest

from metalbm_amd import (Algorithm, Config, ForceType, ForcingSchemeType, LatticeType,
                         LBMError, Precision)
from metalbm_amd.capi import LBM_E_INVALID, LBM_E_STATE, lib

pytestmark = pytest.mark.gpu


def noalloc(a):
    m, e, w = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    rc = lib().lbm_reduce_noalloc(a._h, ctypes.byref(m), ctypes.byref(e), ctypes.byref(w))
    assert rc == 0, lib().lbm_last_error(a._h)
    return m.value, e.value, w.value


def test_reduce_noalloc_matches_reduce():
    cases = [
        # larger first, then smaller: the second call reuses the array the
        # first one filled further
        Config(lattice=LatticeType.D3Q19, nx=64, ny=48, nz=40, forcing_scheme=ForcingSchemeType.Guo,
               force=ForceType.Constant, amplitude=(1e-5, 2e-5, 0), tau=0.6,
               storage=Precision.F32, compute=Precision.F32, shifted=True),
        Config(lattice=LatticeType.D2Q9, nx=36, ny=20, force=ForceType.Kolmogorov,
               amplitude=(1e-4, 0, 0), wavelength=(9.0, 1, 1), tau=0.7),
    ]
    for k, cfg in enumerate(cases):
        print(f"case reduce_noalloc call {k + 1}: {cfg.nx}x{cfg.ny}x{cfg.nz}", flush=True)
        with Algorithm(cfg) as a:
            a.init_taylor_green(1.0, 0.05)
            with pytest.raises(LBMError) as e:
                a.reduce_noalloc()
            assert e.value.code == LBM_E_STATE
            a.iterate(3, store=True)
            got = noalloc(a)
            ref = a.reduce()
            again = noalloc(a)
        assert got == ref, (got, ref)
        assert again == ref
        assert got[0] > 0 and got[1] > 0 and got[2] > 0


def test_reduce_noalloc_rejects_grids_beyond_its_scratch():
    # 2^29 + 2^19 cells of D2Q9 FP16 (19 GB of populations): one reduction
    # block more than the static scratch holds
    cfg = Config(lattice=LatticeType.D2Q9, nx=32768, ny=16400, storage=Precision.F16,
                 compute=Precision.F32, shifted=True)
    assert cfg.cells > 2 ** 29
    print("case reduce_noalloc oversize", flush=True)
    with Algorithm(cfg) as a:
        m = ctypes.c_double()
        rc = lib().lbm_reduce_noalloc(a._h, ctypes.byref(m), None, None)
        assert rc == LBM_E_INVALID
        assert b"2^29" in lib().lbm_last_error(a._h)
