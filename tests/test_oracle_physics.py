"""Pins the oracle's per-cell operators with closed forms and invariants.

Equilibrium (Equilibrium.h:17-33, 60-81, 106-126), forces (Force.h:154-159,
208-215, 262-267), forcing schemes (ForcingScheme.h:99-117, 141-151, 184-197),
BGK collide-and-stream (Collision.h:134-151) and the periodic pull
(Collision.h:146 with the halo fill of Communication.h:134-180 and
Boundary.h:45-102).  Each check is chosen so a plausible slip (a dropped
term, wrong sign, transposed index, wrong tau) makes it fail.
"""
import math

import numpy as np
import pytest

D2Q9, D3Q19, D3Q27 = 3, 9, 10


def cfg_for(o, lat, **kw):
    base = dict(lattice=lat, nx=6, ny=5, nz=4 if lat != D2Q9 else 1, tau=0.7)
    base.update(kw)
    return o.Config(**base)


def ma3_closed(c, w, rho, u):
    """Ma3 equilibrium written independently of the oracle's operation order."""
    cu = c @ u
    u2 = u @ u
    P = (1 + 3 * cu - 1.5 * u2 + 4.5 * cu ** 2 - 4.5 * cu * u2 + 4.5 * cu ** 3
         + 1.125 * u2 ** 2 - 6.75 * cu ** 2 * u2 + 3.375 * cu ** 4)
    return rho * w * P


@pytest.mark.parametrize("lat", [D2Q9, D3Q19, D3Q27])
def test_equilibrium_ma3(oracle, lat):
    cfg = cfg_for(oracle, lat)
    c = oracle.velocities(lat).astype(float)
    w = oracle.weights(lat)
    rng = np.random.default_rng(1)
    for _ in range(20):
        u = rng.uniform(-0.08, 0.08, c.shape[1])
        rho = rng.uniform(0.9, 1.1)
        fe = oracle.equilibrium(cfg, rho, u)
        np.testing.assert_allclose(fe, ma3_closed(c, w, rho, u), rtol=1e-14)
        # density and momentum recovered (TestMoment.cpp:68-95 pattern)
        assert abs(fe.sum() - rho) < 1e-14 * rho * 10 + 1e-3 * np.abs(u).max() ** 4
        np.testing.assert_allclose(fe @ c, rho * u, atol=1e-15 + 1e-2 * np.abs(u).max() ** 3)
    np.testing.assert_array_equal(oracle.equilibrium(cfg, 1.3, np.zeros(c.shape[1])), 1.3 * w)


@pytest.mark.parametrize("lat", [D2Q9, D3Q27])
def test_equilibrium_exact(oracle, lat):
    cfg = cfg_for(oracle, lat, equilibrium=oracle.EQ_EXACT)
    c = oracle.velocities(lat)
    w = oracle.weights(lat)
    rng = np.random.default_rng(2)
    for _ in range(10):
        u = rng.uniform(-0.1, 0.1, c.shape[1])
        rho = rng.uniform(0.9, 1.1)
        fe = oracle.equilibrium(cfg, rho, u)
        s = np.sqrt(1 + 3 * u ** 2)
        B = (2 * u + s) / (1 - u)
        closed = rho * w * np.prod((2 - s) * B ** c, axis=1)
        np.testing.assert_allclose(fe, closed, rtol=1e-14)
        # the entropic equilibrium has exact density and momentum
        assert abs(fe.sum() - rho) < 1e-14
        np.testing.assert_allclose(fe @ c, rho * u, atol=1e-15)


def test_no_exact_for_d3q19(oracle):
    cfg = cfg_for(oracle, D3Q19, equilibrium=oracle.EQ_EXACT)
    f0 = np.ones((19, cfg.cells)) / 19
    with pytest.raises(ValueError):
        oracle.run(cfg, f0, 1)


def test_forces_closed_form(oracle):
    A = 2.5e-5
    lam = 32.0
    k = cfg_for(oracle, D2Q9, force=oracle.FORCE_KOLMOGOROV, amplitude=(A, 7.0, 9.0),
                wavelength=(lam, 3.0, 3.0))
    assert oracle.force_at(k, 3, 0)[0] == 0.0
    assert abs(oracle.force_at(k, 5, 8)[0] - A) < 1e-20          # y = lambda / 4
    assert abs(oracle.force_at(k, 0, 16)[0]) < 1e-19             # y = lambda / 2
    assert abs(oracle.force_at(k, 0, 24)[0] + A) < 1e-20
    assert oracle.force_at(k, 0, 8)[1] == 0.0                    # only F_x is set
    s = cfg_for(oracle, D3Q19, force=oracle.FORCE_SINUSOIDAL, amplitude=(1.0, 2.0, 3.0),
                wavelength=(8.0, 12.0, 16.0))
    F = oracle.force_at(s, 2, 3, 4)
    np.testing.assert_allclose(F, [1.0, 2.0, 3.0], rtol=1e-15)   # quarter wavelengths
    F = oracle.force_at(s, 1, 1, 1)
    np.testing.assert_allclose(F, [math.sin(2 * math.pi / 8), 2 * math.sin(2 * math.pi / 12),
                                   3 * math.sin(2 * math.pi / 16)], rtol=1e-15)
    c = cfg_for(oracle, D3Q27, force=oracle.FORCE_CONSTANT, amplitude=(1e-6, -2e-6, 3e-6))
    np.testing.assert_array_equal(oracle.force_at(c, 4, 1, 2), [1e-6, -2e-6, 3e-6])


def one_cell_step(o, cfg, f_cell):
    """Apply one step to a uniform field; every cell then sees f_cell pulled."""
    N = cfg.cells
    f0 = np.repeat(np.asarray(f_cell)[:, None], N, axis=1)
    r = o.run(cfg, f0, 1)
    return r


@pytest.mark.parametrize("lat", [D2Q9, D3Q19, D3Q27])
@pytest.mark.parametrize("tau", [0.6, 1.0, 1.7])
def test_guo_source_moments(oracle, lat, tau):
    """Sum_i S_i = 0 and sum_i c_i S_i = (1 - 1/(2 tau)) F with the raw u."""
    F = np.array([1e-4, -2e-4, 3e-4])
    cfg = cfg_for(oracle, lat, force=oracle.FORCE_CONSTANT, amplitude=tuple(F), tau=tau)
    c = oracle.velocities(lat).astype(float)
    D = c.shape[1]
    rng = np.random.default_rng(3)
    u = rng.uniform(-0.05, 0.05, D)
    fcell = oracle.equilibrium(cfg, 1.02, u)
    r = one_cell_step(oracle, cfg, fcell)
    # reference without force: same state, forcing scheme None
    cfg0 = cfg_for(oracle, lat, force=oracle.FORCE_CONSTANT, amplitude=tuple(F), tau=tau,
                   forcing_scheme=oracle.FS_NONE)
    r0 = one_cell_step(oracle, cfg0, fcell)
    S = r.f[:, 0] - r0.f[:, 0]
    assert abs(S.sum()) < 1e-15  # S is a difference of O(1) populations
    np.testing.assert_allclose(S @ c, (1 - 1 / (2 * tau)) * F[:D], rtol=1e-10)
    # closed form of each S_i
    cu = c @ u
    S_closed = (1 - 1 / (2 * tau)) * oracle.weights(lat) * 3 * (
        (c - u) @ F[:D] + 3 * cu * (c @ F[:D]))
    np.testing.assert_allclose(S, S_closed, rtol=1e-9, atol=1e-20)


@pytest.mark.parametrize("lat", [D2Q9, D3Q27])
def test_edm_and_shanchen(oracle, lat):
    F = np.array([2e-4, 1e-4, -1e-4])
    c = oracle.velocities(lat).astype(float)
    D = c.shape[1]
    base = dict(force=oracle.FORCE_CONSTANT, amplitude=tuple(F), tau=0.8)
    cfgE = cfg_for(oracle, lat, forcing_scheme=oracle.FS_EDM, **base)
    cfgN = cfg_for(oracle, lat, forcing_scheme=oracle.FS_NONE, **base)
    cfgS = cfg_for(oracle, lat, forcing_scheme=oracle.FS_SHANCHEN, **base)
    u = np.array([0.01, -0.02, 0.015])[:D]
    fcell = oracle.equilibrium(cfgE, 0.97, u)
    rE, rN, rS = (one_cell_step(oracle, k, fcell) for k in (cfgE, cfgN, cfgS))
    S = rE.f[:, 0] - rN.f[:, 0]
    assert abs(S.sum()) < 1e-15
    np.testing.assert_allclose(S @ c, F[:D], rtol=1e-9)           # EDM: full F
    np.testing.assert_array_equal(rS.f, rN.f)                      # Shan-Chen source is 0
    # hydrodynamic velocity: u + F / (2 rho) except for None
    rho = fcell.sum()
    np.testing.assert_allclose(rS.u[:, 0], u + F[:D] / (2 * rho), rtol=1e-12)
    np.testing.assert_allclose(rN.u[:, 0], u, rtol=1e-12)


@pytest.mark.parametrize("lat", [D2Q9, D3Q19, D3Q27])
def test_bgk_invariants(oracle, lat):
    cfg = cfg_for(oracle, lat, force=oracle.FORCE_CONSTANT, amplitude=(1e-5, 2e-5, -1e-5),
                  tau=0.55)
    rng = np.random.default_rng(4)
    u = rng.uniform(-1e-2, 1e-2, (cfg.d, cfg.cells))
    rho = 1 + rng.uniform(-1e-2, 1e-2, cfg.cells)
    f0 = oracle.init_equilibrium(cfg, rho, u)
    c = oracle.velocities(lat).astype(float)
    r = oracle.run(cfg, f0, 3)
    # mass conserved to round-off
    assert abs(r.f.sum() - f0.sum()) < 1e-12 * f0.sum()
    # momentum gains (1 - 1/(2 tau)) F per cell per step (Guo, raw u)
    F = np.array(cfg.amplitude[:cfg.d])
    gain = (c.T @ r.f).sum(axis=1) - (c.T @ f0).sum(axis=1)
    np.testing.assert_allclose(gain, 3 * cfg.cells * (1 - 1 / (2 * cfg.tau)) * F, rtol=1e-7)
    # rest state is a fixed point
    cfg0 = cfg_for(oracle, lat, tau=0.55)
    w = oracle.weights(lat)
    frest = np.repeat(w[:, None], cfg0.cells, axis=1) * 1.1
    np.testing.assert_allclose(oracle.run(cfg0, frest, 5).f, frest, rtol=1e-15)


@pytest.mark.parametrize("lat", [D2Q9, D3Q19, D3Q27])
def test_tau_one_is_equilibrium_plus_source(oracle, lat):
    cfg = cfg_for(oracle, lat, force=oracle.FORCE_CONSTANT, amplitude=(1e-4, 0, 0), tau=1.0)
    rng = np.random.default_rng(5)
    f0 = oracle.init_equilibrium(cfg, 1 + rng.uniform(-1e-2, 1e-2, cfg.cells),
                                 rng.uniform(-1e-2, 1e-2, (cfg.d, cfg.cells)))
    f0 *= 1 + rng.uniform(-1e-3, 1e-3, f0.shape)
    r = oracle.run(cfg, f0, 1)
    # tau = 1: f' = feq(rho, u of the pulled populations) + S
    cfgN = cfg_for(oracle, lat, force=oracle.FORCE_CONSTANT, amplitude=(1e-4, 0, 0),
                   tau=1.0, forcing_scheme=oracle.FS_NONE)
    rN = oracle.run(cfgN, f0, 1)
    for cell in (0, 7, cfg.cells - 1):
        fe = oracle.equilibrium(cfg, rN.rho[cell], rN.u[:, cell])
        np.testing.assert_allclose(rN.f[:, cell], fe, rtol=1e-14)


@pytest.mark.parametrize("lat,shape", [(D2Q9, (5, 7, 1)), (D3Q19, (4, 3, 5)),
                                       (D3Q27, (3, 4, 2))])
def test_pull_is_periodic_roll(oracle, lat, shape):
    """Brute force: with tau -> infinity the step is the pure pull-stream,
    so population i must equal np.roll of itself by c_i."""
    nx, ny, nz = shape
    cfg = cfg_for(oracle, lat, nx=nx, ny=ny, nz=nz, tau=1e15)
    rng = np.random.default_rng(6)
    Q = cfg.q
    f0 = rng.uniform(0.5, 1.5, (Q, cfg.cells)) / Q
    r = oracle.run(cfg, f0, 1)
    c = oracle.velocities(lat)
    for i in range(Q):
        g = f0[i].reshape(nx, ny, nz)
        shift = tuple(int(v) for v in c[i]) + ((0,) if cfg.d == 2 else ())
        expect = np.roll(g, shift, axis=(0, 1, 2)).ravel()
        np.testing.assert_allclose(r.f[i], expect, rtol=1e-13)
    # the stored density is the sum of the pulled populations
    pulled = np.array([np.roll(f0[i].reshape(nx, ny, nz),
                               tuple(int(v) for v in c[i]) + ((0,) if cfg.d == 2 else ()),
                               axis=(0, 1, 2)).ravel() for i in range(Q)])
    np.testing.assert_allclose(r.rho, pulled.sum(axis=0), rtol=1e-15)
