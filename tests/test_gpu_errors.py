"""Error behaviour of live handles: every rejected call returns an LBM_E_*
code with a message and leaves the handle usable (the reference prints and
exits instead, Commons.h:6-38, and fails silently with alpha = 2 in the
solver, Collision.h:344-346)."""
import numpy as np
import pytest

from metalbm_amd import (AlphaMode, Algorithm, CollisionType, Config, LatticeType, LBMError)
from metalbm_amd.capi import LBM_E_HIP, LBM_E_INVALID, LBM_E_NOMEM, LBM_E_STATE

pytestmark = pytest.mark.gpu


def expect(code, needle, fn, *args):
    with pytest.raises(LBMError) as e:
        fn(*args)
    assert e.value.code == code, e.value
    assert needle in e.value.message, e.value.message


def elbm(lat=LatticeType.D2Q9, shape=(16, 16, 1)):
    return Config(lattice=lat, collision=CollisionType.ELBM, nx=shape[0], ny=shape[1],
                  nz=shape[2], tau=0.6)


def test_learned_mode_needs_a_network():
    print("case errors learned-without-network", flush=True)
    with Algorithm(elbm()) as a:
        a.init_noise(1.0, 1e-3, 1)
        a.set_alpha_mode(AlphaMode.LEARNED, 1e-10)
        expect(LBM_E_STATE, "network", a.iterate, 1)
        a.set_alpha_mode(AlphaMode.NEWTON)
        a.iterate(2)  # still usable
        assert a.steps_done == 2


def test_training_needs_samples():
    print("case errors train-without-samples", flush=True)
    with Algorithm(elbm()) as a:
        expect(LBM_E_STATE, "logged samples", a.mlp_train, 10, 8, 1e-3)


def test_learned_alpha_lattices_and_collision():
    print("case errors learned-alpha scope", flush=True)
    with Algorithm(elbm(LatticeType.D3Q19, (4, 4, 4))) as a:
        expect(LBM_E_INVALID, "D2Q9 and D3Q27", a.set_alpha_mode, AlphaMode.LEARNED, 1e-10)
        expect(LBM_E_INVALID, "D2Q9 and D3Q27", a.sample_log_enable, 10, 0.5)
        a.set_alpha_mode(AlphaMode.WARM_RESIDUAL, 1e-10)  # needs no network
        a.init_noise(1.0, 1e-3, 1)
        a.iterate(1)
    with Algorithm(Config(lattice=LatticeType.D2Q9, nx=16, ny=16)) as a:
        expect(LBM_E_STATE, "ELBM", a.set_alpha_mode, AlphaMode.NEWTON)
        expect(LBM_E_STATE, "ELBM", a.mlp_init, 1)


def test_bad_arguments_are_rejected():
    print("case errors bad arguments", flush=True)
    with Algorithm(elbm()) as a:
        expect(LBM_E_INVALID, "nsteps", a.iterate, -1)
        expect(LBM_E_INVALID, "alpha_mode", a.set_alpha_mode, 5)
        expect(LBM_E_INVALID, "residual_tol", a.set_alpha_mode, AlphaMode.WARM_RESIDUAL, 0.0)
        expect(LBM_E_INVALID, "capacity", a.sample_log_enable, 0, 0.5)
        expect(LBM_E_INVALID, "fraction", a.sample_log_enable, 10, 1.5)
        expect(LBM_E_INVALID, "kmax", a.init_spectral, 1.0, 0.01, 8)
        expect(LBM_E_INVALID, "kpeak", a.init_spectral, 1.0, 0.01, 3, 0.0)
        expect(LBM_E_INVALID, "bad samples", a.mlp_predict, np.zeros((0, 9), np.float32))
        a.init_spectral(1.0, 0.01, 3, 2.0)  # the handle still works
        a.iterate(1)


def test_failed_sample_log_enable_logs_nothing():
    """A sample-log enable whose device allocation fails returns an error and
    leaves logging off: the following Newton steps log nothing and the handle
    keeps working; a later feasible enable logs again."""
    print("case errors failed sample-log enable", flush=True)
    cfg = Config(lattice=LatticeType.D2Q9, collision=CollisionType.ELBM, nx=32, ny=32,
                 tau=0.5003)
    with Algorithm(cfg) as a:
        a.init_noise(1.0, 3e-2, 7)
        a.sample_log_enable(4096, 1.0, 1)
        a.iterate(2)
        assert len(a.sample_log()[1]) > 0
        # 2^40 samples of 9 FP32 features: 36 TB, beyond any device
        with pytest.raises(LBMError) as e:
            a.sample_log_enable(2 ** 40, 1.0, 1)
        assert e.value.code in (LBM_E_HIP, LBM_E_NOMEM), e.value
        a.iterate(2)
        assert len(a.sample_log()[1]) == 0
        st = a.elbm_stats()
        assert st["newton"] > 0   # the steps did solve cells, and logged none
        a.sample_log_enable(4096, 1.0, 1)
        a.iterate(1)
        assert len(a.sample_log()[1]) > 0
