"""Pins the spectral-diagnostics oracle (oracle/spectral.py) by closed forms:
curls of single Fourier modes and of the ABC (Beltrami) flow, the reference
scaling's integer wavenumbers, double 1/N and dropped 2D imaginary part
(Transformer.h:100-295), its +n/2 Nyquist derivative off the self-paired
planes, the shell binning floor(|k|) vs round(|k|), the k_last plane
weights, and the spectral initial field's exact shell spectrum."""
import numpy as np
import pytest

from oracle import spectral as sp

TAU = 2.0 * np.pi


def grid(shape):
    return np.meshgrid(*[np.arange(n, dtype=np.float64) for n in shape], indexing="ij")


def field(shape, *comps):
    return np.stack([np.broadcast_to(c, shape).reshape(-1) for c in comps])


def test_2d_physical_curl_of_single_modes():
    shape = (16, 12)
    x, y = grid(shape)
    A = 0.3
    u = field(shape, 0 * x, A * np.sin(TAU * 2 * x / 16))
    w, ens = sp.vorticity(u, shape, sp.PHYSICAL)
    exact = A * (TAU * 2 / 16) * np.cos(TAU * 2 * x / 16)
    np.testing.assert_allclose(w[0], exact.reshape(-1), atol=1e-14)
    assert ens == pytest.approx(0.25 * (A * TAU * 2 / 16) ** 2, rel=1e-13)
    u = field(shape, A * np.sin(TAU * 3 * y / 12), 0 * x)
    w, _ = sp.vorticity(u, shape, sp.PHYSICAL)
    np.testing.assert_allclose(w[0], (-A * (TAU * 3 / 12) * np.cos(TAU * 3 * y / 12)).reshape(-1),
                               atol=1e-14)


def test_3d_physical_curl_abc_flow_is_beltrami():
    n, m = 16, 2
    shape = (n, n, n)
    x, y, z = grid(shape)
    k = TAU * m / n
    A, B, C = 0.3, 0.2, 0.1
    u = field(shape, A * np.sin(k * z) + C * np.cos(k * y),
              B * np.sin(k * x) + A * np.cos(k * z),
              C * np.sin(k * y) + B * np.cos(k * x))
    w, _ = sp.vorticity(u, shape, sp.PHYSICAL)
    np.testing.assert_allclose(w, k * u, atol=1e-14)
    # reference scaling: integer wavenumbers and a second 1/N
    wr, _ = sp.vorticity(u, shape, sp.REFERENCE)
    np.testing.assert_allclose(wr, m * u / n ** 3, atol=1e-17)


def test_3d_curl_component_order_on_a_non_cube():
    shape = (8, 10, 12)
    x, y, z = grid(shape)
    N = 8 * 10 * 12
    u = field(shape, 0 * x, 0 * x, np.sin(TAU * x / 8))
    w, _ = sp.vorticity(u, shape, sp.PHYSICAL)
    # curl (0, 0, uz(x)) = (0, -duz/dx, 0)
    np.testing.assert_allclose(w[1], (-(TAU / 8) * np.cos(TAU * x / 8)).reshape(-1), atol=1e-14)
    np.testing.assert_allclose(w[[0, 2]], 0, atol=1e-14)
    u = field(shape, np.sin(TAU * 2 * y / 10), 0 * x, 0 * x)
    w, _ = sp.vorticity(u, shape, sp.REFERENCE)
    # curl (ux(y), 0, 0) = (0, 0, -dux/dy), integer k = 2, / N^2 overall
    np.testing.assert_allclose(w[2], (-2 * np.cos(TAU * 2 * y / 10) / N).reshape(-1), atol=1e-16)
    np.testing.assert_allclose(w[:2], 0, atol=1e-16)


def test_2d_reference_curl_drops_the_imaginary_part():
    shape = (16, 12)
    x, _ = grid(shape)
    N = 16 * 12
    u = field(shape, 0 * x, np.sin(TAU * 3 * x / 16))  # omega_hat real: kept
    w, _ = sp.vorticity(u, shape, sp.REFERENCE)
    np.testing.assert_allclose(w[0], (3 * np.cos(TAU * 3 * x / 16) / N).reshape(-1), atol=1e-16)
    u = field(shape, 0 * x, np.cos(TAU * 3 * x / 16))  # omega_hat imaginary: dropped
    w, _ = sp.vorticity(u, shape, sp.REFERENCE)
    np.testing.assert_allclose(w[0], 0, atol=1e-16)
    w, _ = sp.vorticity(u, shape, sp.PHYSICAL)
    np.testing.assert_allclose(w[0], (-(TAU * 3 / 16) * np.sin(TAU * 3 * x / 16)).reshape(-1),
                               atol=1e-14)


def test_nyquist_derivative_conventions():
    # (-1)^x on the self-paired plane k_y = 0: zero in both scalings
    shape = (8, 6)
    x, y = grid(shape)
    N = 48
    alt = (-1.0) ** x
    for s in (sp.REFERENCE, sp.PHYSICAL):
        w, _ = sp.vorticity(field(shape, 0 * x, alt), shape, s)
        np.testing.assert_allclose(w[0], 0, atol=1e-15)
    # off the self-paired planes the reference's k_x = +n/2 is used:
    # uy = (-1)^x sin(2 pi y/ny) -> omega = (nx/2) (-1)^x cos(2 pi y/ny) / N
    u = field(shape, 0 * x, alt * np.sin(TAU * y / 6))
    w, _ = sp.vorticity(u, shape, sp.REFERENCE)
    np.testing.assert_allclose(w[0], (4 * alt * np.cos(TAU * y / 6) / N).reshape(-1), atol=1e-15)
    w, _ = sp.vorticity(u, shape, sp.PHYSICAL)
    np.testing.assert_allclose(w[0], 0, atol=1e-15)
    # 3D: uy = (-1)^x sin(2 pi z/nz): omega_x = -duy/dz, omega_z = duy/dx
    shape = (6, 4, 8)
    x, y, z = grid(shape)
    N = 6 * 4 * 8
    alt = (-1.0) ** x
    u = field(shape, 0 * x, alt * np.sin(TAU * z / 8), 0 * x)
    w, _ = sp.vorticity(u, shape, sp.REFERENCE)
    np.testing.assert_allclose(w[0], (-alt * np.cos(TAU * z / 8) / N).reshape(-1), atol=1e-15)
    np.testing.assert_allclose(w[2], (3 * alt * np.cos(TAU * z / 8) / N).reshape(-1), atol=1e-15)
    w, _ = sp.vorticity(u, shape, sp.PHYSICAL)
    np.testing.assert_allclose(w[0], (-(TAU / 8) * alt * np.cos(TAU * z / 8)).reshape(-1),
                               atol=1e-14)
    np.testing.assert_allclose(w[2], 0, atol=1e-14)


def test_energy_spectrum_single_modes_and_plane_weights():
    shape = (16, 12)
    x, y = grid(shape)
    N = 16 * 12
    A = 0.25
    ns = sp.shells(shape)
    assert ns == 8
    for u in (field(shape, A * np.sin(TAU * 2 * x / 16), 0 * x),   # k_y = 0 plane, c = 1/2 twice
              field(shape, 0 * x, A * np.sin(TAU * 2 * y / 12))):  # interior k_y, c = 1
        e = sp.spectrum(u, shape, sp.REFERENCE)
        expect = np.zeros(ns)
        expect[2] = A * A * N / 4
        np.testing.assert_allclose(e, expect, atol=1e-13)
        f = sp.spectrum(u, shape, sp.REFERENCE, forcing=True)
        np.testing.assert_allclose(f, expect * N, atol=1e-10)  # forcing: not normalised
        p = sp.spectrum(u, shape, sp.PHYSICAL)
        np.testing.assert_allclose(p, expect / N, atol=1e-16)  # = 1/2 mean u^2 = A^2/4


def test_shell_binning_floor_vs_round():
    shape = (16, 16)
    x, y = grid(shape)
    u = field(shape, np.sin(TAU * (2 * x + 2 * y) / 16), 0 * x)  # |k| = 2.83
    e = sp.spectrum(u, shape, sp.REFERENCE)
    p = sp.spectrum(u, shape, sp.PHYSICAL)
    assert np.flatnonzero(np.abs(e) > 1e-12).tolist() == [2]  # int(2.83)
    assert np.flatnonzero(np.abs(p) > 1e-16).tolist() == [3]  # round(2.83)
    assert p[3] == pytest.approx(0.25, rel=1e-13)


def test_nyquist_plane_weight():
    shape = (16, 8)
    x, y = grid(shape)
    A = 0.5
    u = field(shape, 0 * x, A * (-1.0) ** y)  # the k_y = n/2 plane only
    p = sp.spectrum(u, shape, sp.PHYSICAL)
    assert p[4] == pytest.approx(A * A / 2, rel=1e-13)  # = 1/2 mean |u|^2
    e = sp.spectrum(u, shape, sp.REFERENCE)
    assert e[4] == pytest.approx(A * A * 128, rel=1e-13)  # reference weight 1 there


@pytest.mark.parametrize("lattice,shape", [(3, (24, 20, 1)), (10, (12, 10, 14))])
def test_spectral_init_has_its_prescribed_shell_spectrum(oracle, lattice, shape):
    cfg = oracle.Config(lattice=lattice, nx=shape[0], ny=shape[1], nz=shape[2])
    urms, kmax, kpeak = 0.05, 4, 2.0
    u = oracle.spectral_velocity(cfg, urms, kmax, kpeak, 3)
    D = u.shape[0]
    sh = shape[:D]
    p = sp.spectrum(u, sh, sp.PHYSICAL)
    s = np.arange(1, kmax + 1)
    E = s ** 4.0 * np.exp(-2.0 * (s / kpeak) ** 2)
    gamma2 = urms ** 2 * D / E.sum()
    expect = np.zeros(len(p))
    expect[1:kmax + 1] = 0.5 * gamma2 * E
    np.testing.assert_allclose(p, expect, rtol=1e-12, atol=1e-18)
    assert p.sum() == pytest.approx(0.5 * np.mean(np.sum(u * u, axis=0)), rel=1e-12)
