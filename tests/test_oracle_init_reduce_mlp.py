"""Pins the oracle's initial conditions, diagnostics and MLP reference.

* SplitMix64: the published first outputs of the stream seeded with 1234567
  (Vigna's reference implementation, used by xoshiro seeding).
* Noise: range and draw order (cell-major, component-minor).
* Taylor-Green: closed-form samples and E(0) = rho0 U0^2 / 8 (3D) on a node
  grid, U0^2 / 4 (2D).
* ref_reduce: uniform flow E = 1/2 rho |u|^2, zero curl; TGV enstrophy against
  the second-order central-difference closed form.
* MLP gradient: central finite differences of the FP64 loss.
* Adam: one step restated from Kingma & Ba (2015).
"""
import math

import numpy as np
import pytest

D2Q9, D3Q19, D3Q27 = 3, 9, 10


def test_splitmix64_published_vectors(oracle):
    expect = [6457827717110365317, 3203168211198807973, 9817491932198370423,
              4593380528125082431, 16408922859458223821]
    assert [oracle.splitmix64(1234567, k) for k in range(5)] == expect


def test_noise_order_and_range(oracle):
    cfg = oracle.Config(lattice=D3Q19, nx=4, ny=3, nz=5)
    u = oracle.noise_velocity(cfg, 42, 1e-3)
    assert u.shape == (3, 60)
    assert np.all(np.abs(u) <= 1e-3)
    for cell, d in [(0, 0), (7, 2), (59, 1)]:
        r = oracle.splitmix64(42, cell * 3 + d)
        assert u[d, cell] == 1e-3 * (2.0 * ((r >> 11) * 2.0 ** -53) - 1.0)
    big = oracle.noise_velocity(oracle.Config(lattice=D2Q9, nx=64, ny=64), 7, 1.0)
    assert abs(big.mean()) < 0.02 and abs(big.std() - 1 / math.sqrt(3)) < 0.01


@pytest.mark.parametrize("lat,shape", [(D2Q9, (8, 12, 1)), (D3Q27, (8, 6, 4))])
def test_taylor_green(oracle, lat, shape):
    nx, ny, nz = shape
    cfg = oracle.Config(lattice=lat, nx=nx, ny=ny, nz=nz)
    U0 = 0.05
    u = oracle.taylor_green(cfg, U0)
    x, y, z = 3, 5, 1 if lat != D2Q9 else 0
    cell = (x * ny + y) * nz + z
    X, Y, Z = 2 * math.pi * x / nx, 2 * math.pi * y / ny, 2 * math.pi * z / nz
    cz = math.cos(Z) if lat != D2Q9 else 1.0
    assert abs(u[0, cell] - U0 * math.sin(X) * math.cos(Y) * cz) < 1e-17
    assert abs(u[1, cell] + U0 * math.cos(X) * math.sin(Y) * cz) < 1e-17
    rho = np.ones(cfg.cells)
    m, e, w = oracle.reduce(cfg, rho, u)
    assert abs(m - cfg.cells) < 1e-12 * cfg.cells
    expect_e = U0 ** 2 / 8 if lat != D2Q9 else U0 ** 2 / 4
    assert abs(e - expect_e) < 1e-15
    # enstrophy of the discrete curl: central differences scale each
    # derivative of sin/cos(k x) by sin(k)/k with k = 2 pi / n.
    kx, ky, kz = (math.sin(2 * math.pi / n) for n in (nx, ny, nz))
    if lat == D2Q9:
        # w = -U0 (kx + ky) sin x sin y ; <sin^2 sin^2> = 1/4
        expect_w = 0.5 * U0 ** 2 * (kx + ky) ** 2 / 4
    else:
        # wx = -U0 kz cos x sin y sin z, wy = -U0 kz sin x cos y sin z,
        # wz = U0 (kx + ky) sin x sin y cos z ; each mean square is 1/8
        expect_w = 0.5 * U0 ** 2 * (2 * kz ** 2 + (kx + ky) ** 2) / 8
    assert abs(w - expect_w) < 1e-14 * max(1.0, expect_w) + 1e-18


def test_reduce_uniform_flow(oracle):
    cfg = oracle.Config(lattice=D3Q19, nx=5, ny=4, nz=3)
    rho = np.full(cfg.cells, 1.2)
    u = np.tile(np.array([[0.01], [-0.02], [0.03]]), (1, cfg.cells))
    m, e, w = oracle.reduce(cfg, rho, u)
    assert abs(m - 1.2 * 60) < 1e-12
    assert abs(e - 0.5 * 1.2 * (0.01 ** 2 + 0.02 ** 2 + 0.03 ** 2)) < 1e-17
    assert w == 0.0


@pytest.mark.parametrize("dims", [(9, 32, 32), (27, 16, 8)])
def test_mlp_gradient_finite_differences(oracle, dims):
    nin, h1, h2 = dims
    n = oracle.mlp_param_count(nin, h1, h2)
    assert n == h1 * nin + h1 + h2 * h1 + h2 + h2 + 1
    rng = np.random.default_rng(8)
    p = rng.normal(0, 0.4, n)
    x = rng.normal(0, 1, (13, nin))
    t = rng.normal(0, 1, 13)
    loss, g = oracle.mlp_grad(p, nin, h1, h2, x, t)
    assert abs(loss - oracle.mlp_loss(p, nin, h1, h2, x, t)) < 1e-15
    y = oracle.mlp_forward(p, nin, h1, h2, x)
    assert abs(loss - 0.5 * np.mean((y - t) ** 2)) < 1e-14
    idx = rng.choice(n, 40, replace=False)
    for k in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (oracle.mlp_loss(pp, nin, h1, h2, x, t) - oracle.mlp_loss(pm, nin, h1, h2, x, t)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-9


def test_adam_step(oracle):
    rng = np.random.default_rng(9)
    p = rng.normal(size=7)
    g = rng.normal(size=7)
    m = rng.normal(size=7) * 0.1
    v = np.abs(rng.normal(size=7)) * 0.1
    p0, m0, v0 = p.copy(), m.copy(), v.copy()
    oracle.mlp_adam(p, g, m, v, 1e-3, 0.9, 0.999, 1e-8, 3)
    me = 0.9 * m0 + 0.1 * g
    ve = 0.999 * v0 + 0.001 * g * g
    pe = p0 - 1e-3 * (me / (1 - 0.9 ** 3)) / (np.sqrt(ve / (1 - 0.999 ** 3)) + 1e-8)
    np.testing.assert_allclose(m, me, rtol=1e-15)
    np.testing.assert_allclose(v, ve, rtol=1e-15)
    np.testing.assert_allclose(p, pe, rtol=1e-15)
