"""Spectral diagnostics on the GPU (hipFFT transforms + the library's curl and
shell kernels) against the numpy oracle on the same stored fields.

The two FFT implementations round differently, so the bar is relative to the
largest magnitude: 1e-12 for the vorticity, 1e-11 for the shell sums."""
import numpy as np
import pytest

from metalbm_amd import (Algorithm, CollisionType, Config, ForceType, LatticeType, LBMError,
                         Precision)
from metalbm_amd.capi import LBM_E_INVALID, LBM_E_STATE
from oracle import spectral as sp

pytestmark = pytest.mark.gpu

CASES = [
    # name, config, init, steps
    ("d2q9-kolmogorov-64x48", Config(lattice=LatticeType.D2Q9, nx=64, ny=48,
                                     force=ForceType.Kolmogorov, amplitude=(1e-4, 0, 0),
                                     wavelength=(16.0, 1, 1)), "noise", 20),
    ("d2q9-odd-45x33", Config(lattice=LatticeType.D2Q9, nx=45, ny=33,
                              force=ForceType.Sinusoidal, amplitude=(1e-4, 2e-4, 0),
                              wavelength=(15.0, 11.0, 1)), "noise", 10),
    ("d3q19-spectral-32x24x20", Config(lattice=LatticeType.D3Q19, nx=32, ny=24, nz=20,
                                       force=ForceType.Constant, amplitude=(1e-5, 0, 2e-5)),
     "spectral", 5),
    ("d3q27-elbm-odd-16x18x15", Config(lattice=LatticeType.D3Q27, collision=CollisionType.ELBM,
                                       nx=16, ny=18, nz=15, tau=0.5003), "noise", 5),
    # 2200 shells: two shell windows of the binning kernel
    ("d2q9-wide-4400x8", Config(lattice=LatticeType.D2Q9, nx=4400, ny=8,
                                force=ForceType.Kolmogorov, amplitude=(1e-4, 0, 0),
                                wavelength=(8.0, 1, 1)), "noise", 3),
    ("d3q19-f32-64", Config(lattice=LatticeType.D3Q19, nx=64, ny=64, nz=64,
                            storage=Precision.F32, compute=Precision.F32, shifted=True,
                            force=ForceType.Kolmogorov, amplitude=(1e-5, 0, 0),
                            wavelength=(32.0, 1, 1)), "spectral", 5),
]


def run(c, init, steps):
    a = Algorithm(c)
    if init == "noise":
        a.init_noise(1.0, 2e-2, 11)
    else:
        a.init_spectral(1.0, 0.02, 3, 2.0, 5)
    a.iterate(steps, store=True)
    return a


def rel_err(got, ref):
    scale = max(np.max(np.abs(ref)), 1e-300)
    return float(np.max(np.abs(got - ref)) / scale)


@pytest.mark.parametrize("name,c,init,steps", CASES, ids=[x[0] for x in CASES])
def test_vorticity_and_spectra_match_oracle(name, c, init, steps):
    print(f"case spectral {name}", flush=True)
    with run(c, init, steps) as a:
        rho, u, al, F = a.store_fields()
        shape = (c.nx, c.ny) if c.lattice == LatticeType.D2Q9 else (c.nx, c.ny, c.nz)
        assert a.spectral_shells() == sp.shells(shape)
        for scaling in (sp.REFERENCE, sp.PHYSICAL):
            w, ens = a.vorticity(scaling)
            wo, enso = sp.vorticity(u, shape, scaling)
            assert rel_err(w, wo) <= 1e-12, (scaling, rel_err(w, wo))
            assert ens == pytest.approx(enso, rel=1e-11)
            e, f = a.spectra(scaling)
            eo = sp.spectrum(u, shape, scaling)
            fo = sp.spectrum(F, shape, scaling, forcing=True)
            assert rel_err(e, eo) <= 1e-11, (scaling, rel_err(e, eo))
            if np.any(fo):
                assert rel_err(f, fo) <= 1e-11, (scaling, rel_err(f, fo))
            else:
                assert not np.any(f)
            # deterministic: a second call gives the same bits
            w2, ens2 = a.vorticity(scaling)
            e2, f2 = a.spectra(scaling)
            assert np.array_equal(w, w2) and ens == ens2
            assert np.array_equal(e, e2) and np.array_equal(f, f2)
        # physical energy spectrum: Parseval below the last shell
        e, _ = a.spectra("physical")
        mean_half_u2 = 0.5 * np.mean(np.sum(u * u, axis=0))
        assert e.sum() <= mean_half_u2 * (1 + 1e-12)


def test_fields_and_scaling_errors():
    print("case spectral errors", flush=True)
    with Algorithm(Config(lattice=LatticeType.D2Q9, nx=16, ny=16)) as a:
        a.init_noise(1.0, 1e-2, 1)
        with pytest.raises(LBMError) as e:
            a.vorticity("physical")
        assert e.value.code == LBM_E_STATE
        a.iterate(1, store=True)
        from metalbm_amd.capi import lib
        import ctypes
        rc = lib().lbm_spectra(a._h, 7, None, None)
        assert rc == LBM_E_INVALID
        a.vorticity("physical")  # still usable
