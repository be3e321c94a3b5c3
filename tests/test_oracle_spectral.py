"""The oracle's spectral (solenoidal random-phase) initialiser, pinned by
properties that an FFT of its output must show and by hand-derived modes.

numpy's FFT recovers the mode amplitudes: fftn(u)/N = gamma * uh(k) for the
convention u(x) = sum_k uh(k) exp(+2 pi i k.x/n).  From them the tests check
* support: zero outside the active shells 1..kmax;
* solenoidality: k . uh(k) = 0 for every mode;
* spectrum: |uh(k)|^2 = gamma^2 E(s)/N_s for every active mode, with N_s
  counted here by brute force;
* normalisation: mean |u|^2 = D urms^2 (Parseval);
* draw order, polarisation basis and sign convention, through the first
  half-space mode worked out by hand: k = (0, 0, 1) in 3D, (0, 1) in 2D.
"""
import numpy as np
import pytest

E = lambda s, kp: s ** 4 * np.exp(-2.0 * (s / kp) ** 2)


def draws(o, seed, j):
    return 2 * np.pi * (o.splitmix64(seed, j) >> 11) * 2.0 ** -53


def modes(u, shape):
    N = int(np.prod(shape))
    return [np.fft.fftn(u[d].reshape(shape)) / N for d in range(u.shape[0])]


def wavevectors(shape):
    ks = np.meshgrid(*[np.fft.fftfreq(n, 1.0 / n).round().astype(int) for n in shape],
                     indexing="ij")
    return ks


@pytest.mark.parametrize("d,shape,kmax,kpeak", [(3, (24, 20, 18), 6, 3.0),
                                                (2, (22, 26), 5, 2.5)])
def test_spectral_init_properties(oracle, d, shape, kmax, kpeak):
    lat = oracle.D3Q27 if d == 3 else oracle.D2Q9
    full = shape if d == 3 else shape + (1,)
    cfg = oracle.Config(lattice=lat, nx=full[0], ny=full[1], nz=full[2])
    urms, seed = 0.03, 42
    u = oracle.spectral_velocity(cfg, urms, kmax, kpeak, seed)
    assert u.shape == (d, int(np.prod(shape)))
    # normalisation (Parseval)
    assert abs(np.mean(np.sum(u * u, axis=0)) - d * urms ** 2) <= 1e-12 * d * urms ** 2
    U = modes(u, shape)
    ks = wavevectors(shape)
    k2 = sum(k * k for k in ks)
    s = np.rint(np.sqrt(k2)).astype(int)
    inbox = np.all([np.abs(k) <= kmax for k in ks], axis=0)
    active = inbox & (s >= 1) & (s <= kmax)
    power = sum(np.abs(c) ** 2 for c in U)
    scale = np.sqrt(sum(np.abs(c) ** 2 for c in U).max())
    # support
    assert power[~active].max() <= (1e-12 * scale) ** 2
    # solenoidal
    div = sum(k * c for k, c in zip(ks, U))
    assert np.abs(div).max() <= 1e-12 * kmax * scale
    # spectrum, mode by mode: |uh|^2 N_s / E(s) is the same constant gamma^2
    counts = {t: int(np.sum(active & (s == t))) for t in range(1, kmax + 1)}
    esum = sum(E(t, kpeak) for t in range(1, kmax + 1))
    gamma2 = urms ** 2 * d / esum
    ratio = power[active] * np.array([counts[t] for t in s[active]]) / E(s[active], kpeak)
    np.testing.assert_allclose(ratio, gamma2, rtol=1e-10)
    # the first half-space mode, by hand
    t1, t2, ph = (draws(oracle, seed, j) for j in range(3))
    A1 = np.sqrt(E(1, kpeak) / counts[1]) * np.sqrt(gamma2)
    if d == 3:
        # k = (0,0,1): axis x (smallest |k_d|, lowest index), e1 = k x x = (0,1,0),
        # e2 = k x e1 = (-1,0,0)
        want = [-A1 * np.sin(ph) * np.exp(1j * t2), A1 * np.cos(ph) * np.exp(1j * t1), 0.0]
        got = [c[0, 0, 1] for c in U]
    else:
        # k = (0,1): e1 = (-k_y, k_x)/|k| = (-1, 0)
        want = [-A1 * np.exp(1j * t1), 0.0]
        got = [c[0, 1] for c in U]
    np.testing.assert_allclose(np.array(got), np.array(want, dtype=complex), rtol=0,
                               atol=1e-13 * A1)
    # determinism and seed dependence
    np.testing.assert_array_equal(u, oracle.spectral_velocity(cfg, urms, kmax, kpeak, seed))
    assert np.abs(u - oracle.spectral_velocity(cfg, urms, kmax, kpeak, seed + 1)).max() > 1e-3 * urms


def test_spectral_init_rejects_aliasing(oracle):
    cfg = oracle.Config(lattice=oracle.D3Q27, nx=16, ny=16, nz=16)
    with pytest.raises(ValueError):
        oracle.spectral_velocity(cfg, 0.01, 8, 4.0, 1)
    with pytest.raises(ValueError):
        oracle.spectral_velocity(cfg, 0.01, 0, 4.0, 1)
