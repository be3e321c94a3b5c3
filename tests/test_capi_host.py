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
    assert L.lbm_abi_version() == 3
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


# (lattice, extents, storage, compute, shifted, kernel request) -> (shape, V)
# as lbm_create chooses them; 1 = one cell per lane, 2 = vector kernel.
SHAPES = [
    ((9, (256, 256, 256), 1, 1, 1, 0), (2, 4)),     # headline D3Q19 FP32: V = 4
    ((9, (1024, 1024, 1024), 1, 1, 1, 0), (2, 4)),  # 2^32 bytes per population array
    ((9, (1024, 1024, 512), 0, 1, 1, 0), (2, 2)),   # FP64 storage, FP32 compute, 2^32 B
    ((9, (256, 256, 256), 0, 0, 0, 0), (1, 0)),     # FP64 reference order: cell kernel
    ((10, (384, 384, 384), 2, 1, 1, 0), (1, 0)),    # D3Q27 FP16: cell kernel
    ((10, (384, 384, 384), 2, 1, 1, 2), (2, 2)),    # FP16 asked for: V = 2, full segments
    ((10, (384, 384, 384), 1, 1, 1, 2), (2, 4)),    # FP32 asked for: 16-byte lanes
    ((3, (128, 128, 1), 0, 0, 0, 0), (1, 0)),       # C1 parity config
    ((3, (2048, 2048, 1), 2, 1, 1, 2), (2, 8)),     # 2D FP16: V = 8
    ((9, (4, 4, 100), 0, 0, 0, 2), (2, 2)),         # partial segment: 50 of 64 lanes
    ((9, (4, 4, 99), 0, 0, 0, 1), (1, 0)),          # cell kernel asked for
]


@pytest.mark.parametrize("case,expect", SHAPES)
def test_kernel_shape_selection(capi, case, expect):
    """The kernel shape and cells per lane lbm_create would choose, resolved
    on the host (lbm_kernel_shape) -- including grids whose population arrays
    hold 2^32 bytes or more, which the vector kernel addresses through
    per-row buffer bases."""
    from metalbm_amd import Config, kernel_shape
    lat, (nx, ny, nz), storage, compute, shifted, kernel = case
    cfg = Config(lattice=lat, nx=nx, ny=ny, nz=nz, storage=storage, compute=compute,
                 shifted=bool(shifted), kernel=kernel)
    shape, vec = kernel_shape(cfg)
    assert (int(shape), vec) == expect


def test_kernel_shape_rejects_what_create_rejects(capi):
    from metalbm_amd import Config, LBMError, kernel_shape
    with pytest.raises(LBMError, match="divisible by V"):
        kernel_shape(Config(lattice=9, nx=2, ny=2, nz=99, kernel=2))
    with pytest.raises(LBMError, match="2\\^31"):
        kernel_shape(Config(lattice=9, nx=2048, ny=2048, nz=1024))
    with pytest.raises(LBMError, match="BGK only"):
        kernel_shape(Config(lattice=9, nx=4, ny=4, nz=4, collision=2, kernel=2))
