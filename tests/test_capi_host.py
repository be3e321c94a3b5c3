"""Host-only checks of the C ABI: the library loads, exports every entry point
declared in include/metalbm_amd.h, and rejects invalid configurations with
LBM_E_INVALID and a message before touching a GPU (the reference exits
instead, Commons.h:6-38)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "metalbm_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def capi():
    from metalbm_amd import capi
    return capi


def test_every_declared_symbol_is_exported(capi):
    names = declared_functions()
    assert len(names) >= 20
    L = capi.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(capi.EXPORTS)


def test_abi_version_and_defaults(capi):
    L = capi.lib()
    assert L.lbm_abi_version() == 2
    c = capi.LbmConfigC()
    assert L.lbm_config_default(ctypes.byref(c)) == 0
    assert (c.lattice, c.collision, c.forcing_scheme, c.equilibrium) == (3, 1, 2, 3)
    assert (c.nx, c.ny, c.nz, c.storage, c.compute, c.shifted) == (16, 16, 1, 0, 0, 0)
    assert c.tau == 0.6 and c.residual_tol == 1e-10
    assert (c.force_kmin, c.force_kmax, c.removal_kmin, c.removal_kmax) == (1, 2, 1, 2)
    assert list(c.removal_amplitude) == [1e-6] * 3


BAD = [
    (dict(lattice=4), "lattice"),
    (dict(collision=7), "collision"),
    (dict(forcing_scheme=0), "forcing_scheme"),
    (dict(force=1), "force"),
    (dict(force=10), "force"),
    (dict(force=7, force_kmin=3, force_kmax=2), "force_kmin"),
    (dict(force=9, removal_kmin=-1), "removal_kmin"),
    (dict(lattice=9, nz=4, equilibrium=1), "Exact"),
    (dict(nz=2), "nz must be 1"),
    (dict(nx=0), "positive"),
    (dict(tau=0.0), "tau"),
    (dict(force=5, wavelength=(0.0, 1.0, 1.0)), "wavelength"),
    (dict(storage=2, compute=1, shifted=0), "F16"),
    (dict(storage=1, compute=0), "F64 compute"),
    (dict(nx=2048, ny=2048, lattice=9, nz=1024), "2^31"),
]


@pytest.mark.parametrize("kw,needle", BAD)
def test_invalid_configs_are_rejected(capi, kw, needle):
    from metalbm_amd import Config
    cfg = Config(**kw)
    c = cfg.to_c()
    h = ctypes.c_void_p()
    rc = capi.lib().lbm_create(ctypes.byref(c), ctypes.byref(h))
    assert rc == capi.LBM_E_INVALID
    assert not h.value
    msg = capi.lib().lbm_last_error(None).decode()
    assert needle in msg, msg


def test_null_handle_calls_fail_cleanly(capi):
    L = capi.lib()
    assert L.lbm_step(None, 1, 0) == capi.LBM_E_INVALID
    assert L.lbm_synchronize(None) == capi.LBM_E_INVALID
    assert L.lbm_destroy(None) == 0
    assert L.lbm_steps_done(None) == -1
    assert L.lbm_get_config(None, None) == capi.LBM_E_INVALID
    assert L.lbm_write_fields_hdf5(None, b"x.h5", None) == capi.LBM_E_INVALID
    assert L.lbm_write_distribution_hdf5(None, b"x.h5") == capi.LBM_E_INVALID
    assert L.lbm_read_distribution_hdf5(None, b"x.h5") == capi.LBM_E_INVALID
