"""Pins the spectral-force oracle (oracle/spectral.py, Force.h:294-623) by
closed forms: the ConstantShell field of the unit shell in 2D and 3D
(component order, signs, the integer-wavenumber curl and the 1/N), the
EnergyRemoval shell filter on a two-mode momentum, Turbulent2D as their sum,
and the force-field step of the C oracle against its analytic Constant force."""
import numpy as np
import pytest

from oracle import spectral as sp

TAU = 2.0 * np.pi


def grid(shape):
    return np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")


def test_constant_shell_2d_unit_shell():
    # psi = (2A/N)(cos tx + cos ty); F = (dpsi/dy, -dpsi/dx) with integer k
    n, A = 16, 0.3
    N = n * n
    x, y = grid((n, n))
    F = sp.constant_shell_force((n, n), (A, 7.0, 7.0), 1, 1)
    np.testing.assert_allclose(F[0], (-(2 * A / N) * np.sin(TAU * y / n)).ravel(), atol=1e-17)
    np.testing.assert_allclose(F[1], ((2 * A / N) * np.sin(TAU * x / n)).ravel(), atol=1e-17)


def test_constant_shell_3d_unit_shell():
    n = 12
    N = n ** 3
    A0, A1, A2 = 0.3, -0.2, 0.5
    x, y, z = grid((n, n, n))
    sx, sy, sz = (np.sin(TAU * v / n) for v in (x, y, z))
    F = sp.constant_shell_force((n, n, n), (A0, A1, A2), 1, 1)
    c = 2.0 / N
    np.testing.assert_allclose(F[0], (c * (-A2 * sy + A1 * sz)).ravel(), atol=1e-17)
    np.testing.assert_allclose(F[1], (c * (-A0 * sz + A2 * sx)).ravel(), atol=1e-17)
    np.testing.assert_allclose(F[2], (c * (-A1 * sx + A0 * sy)).ravel(), atol=1e-17)


def test_constant_shell_is_solenoidal_and_shell_limited():
    shape = (24, 20)
    F = sp.constant_shell_force(shape, (1e-3, 0, 0), 3, 5)
    Fh = [np.fft.rfftn(F[d].reshape(shape)) for d in range(2)]
    kx = np.fft.fftfreq(24, 1 / 24)[:, None]
    ky = np.arange(20 // 2 + 1)[None, :]
    k2 = kx ** 2 + ky ** 2
    div = kx * Fh[0] + ky * Fh[1]
    assert np.max(np.abs(div)) < 1e-15
    outside = (k2 < 9) | (k2 > 25)
    assert np.max(np.abs(Fh[0][outside])) < 1e-15 and np.max(np.abs(Fh[1][outside])) < 1e-15
    assert np.max(np.abs(Fh[0][~outside])) > 1e-6


def test_energy_removal_filters_the_momentum_shell():
    shape = (16, 20)
    x, y = grid(shape)
    rho = 1.0 + 0.1 * np.cos(TAU * 3 * y / 20)          # density outside the shell
    p_in = 0.02 * np.cos(TAU * x / 16)                  # |k| = 1: in the shell 1..2
    p_out = 0.03 * np.sin(TAU * 5 * y / 20)             # |k| = 5: outside
    u = np.stack([((p_in + p_out) / rho).ravel(), (p_in / rho).ravel()])
    R = (1e-3, 2e-3, 0)
    F = sp.energy_removal_force(rho.ravel(), u, shape, R, 1, 2)
    np.testing.assert_allclose(F[0], -R[0] * p_in.ravel(), atol=1e-17)
    np.testing.assert_allclose(F[1], -R[1] * p_in.ravel(), atol=1e-17)


def test_turbulent_force_is_the_sum():
    shape = (8, 10, 6)
    rng = np.random.default_rng(3)
    rho = 1 + 0.01 * rng.standard_normal(480)
    u = 0.01 * rng.standard_normal((3, 480))
    A, R = (1e-3, 2e-3, 3e-3), (1e-4, 1e-4, 2e-4)
    T = sp.turbulent_force(rho, u, shape, A, 1, 2, R, 1, 3)
    np.testing.assert_allclose(T, sp.constant_shell_force(shape, A, 1, 2)
                               + sp.energy_removal_force(rho, u, shape, R, 1, 3), atol=1e-18)


def test_force_field_step_equals_the_analytic_force(oracle):
    # a Constant force supplied as a field gives the same step bit for bit
    cfg = oracle.Config(lattice=oracle.D2Q9, nx=8, ny=6, force=oracle.FORCE_CONSTANT,
                        amplitude=(1e-3, -2e-3, 0))
    u0 = oracle.noise_velocity(cfg, 5, 2e-2)
    f0 = oracle.init_equilibrium(cfg, np.ones(cfg.cells), u0)
    ref = oracle.run(cfg, f0, 1)
    field = np.stack([np.full(cfg.cells, 1e-3), np.full(cfg.cells, -2e-3)])
    cfg7 = oracle.Config(**{**cfg.__dict__, "force": oracle.FORCE_CONSTANT_SHELL})
    f1, rho, u, F = oracle.step_field(cfg7, f0, None, field)
    assert np.array_equal(f1, ref.f) and np.array_equal(u, ref.u)
    assert np.array_equal(F, field)
    with pytest.raises(ValueError):
        oracle.run(cfg7, f0, 1)  # a spectral force needs its field
