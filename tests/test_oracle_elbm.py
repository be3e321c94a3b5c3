"""Pins the oracle's entropic collision (Collision.h:227-375,
EntropicStep.h:31-62, 111-140) with hand-built cells that force every branch,
an extended-precision bisection for the Newton root, and the BGK limit.
"""
import numpy as np
import pytest

D2Q9, D3Q19, D3Q27 = 3, 9, 10


def cell_state(o, lat, rho, u, rel, seed):
    """Pulled populations f and fneq = f - feq(rho_f, u_f) for a perturbed cell."""
    cfg = o.Config(lattice=lat, nx=2, ny=2, nz=1 if lat == D2Q9 else 2)
    c = o.velocities(lat).astype(float)
    feq0 = o.equilibrium(cfg, rho, u)
    rng = np.random.default_rng(seed)
    # perturb in the non-hydrodynamic subspace so rho and u stay put
    p = rng.uniform(-1, 1, feq0.size) * rel * feq0
    M = np.vstack([np.ones(feq0.size), c.T])
    p -= M.T @ np.linalg.solve(M @ M.T, M @ p)
    f = feq0 + p
    r = f.sum()
    v = (c.T @ f) / r
    fneq = f - o.equilibrium(cfg, r, v)
    return f, fneq


def g_longdouble(lat, w, f, fneq, a):
    f = f.astype(np.longdouble)
    n = fneq.astype(np.longdouble)
    w = w.astype(np.longdouble)
    fm = f - np.longdouble(a) * n
    return np.sum(f * np.log(f / w) - fm * np.log(fm / w))


def bisect_root(lat, w, f, fneq, lo, hi):
    glo = g_longdouble(lat, w, f, fneq, lo)
    for _ in range(200):
        mid = (lo + hi) / 2
        gm = g_longdouble(lat, w, f, fneq, mid)
        if (gm > 0) == (glo > 0):
            lo, glo = mid, gm
        else:
            hi = mid
    return float((lo + hi) / 2)


@pytest.mark.parametrize("lat", [D2Q9, D3Q19, D3Q27])
def test_small_deviation_branch(oracle, lat):
    f, fneq = cell_state(oracle, lat, 1.0, np.full(3 if lat != D2Q9 else 2, 0.01)[: (2 if lat == D2Q9 else 3)], 1e-4, 1)
    assert np.max(np.abs(fneq) / f) <= 1e-3
    a, b, it = oracle.elbm_alpha(lat, f, fneq, 1.7)
    assert (a, b, it) == (2.0, oracle.ALPHA_SMALL, 0)


@pytest.mark.parametrize("lat", [D2Q9, D3Q27])
def test_capped_branch(oracle, lat):
    f, fneq = cell_state(oracle, lat, 1.0, np.zeros(2 if lat == D2Q9 else 3), 0.02, 2)
    fneq = fneq.copy()
    fneq[3] = 0.6 * f[3]            # alphaMax = f/fneq = 1/0.6 < 2
    a, b, it = oracle.elbm_alpha(lat, f, fneq, 2.0)
    assert b == oracle.ALPHA_CAPPED and it == 0
    amax = min(2.5, min(abs(f[i]) / fneq[i] for i in range(f.size) if fneq[i] > 0))
    assert a == 0.95 * amax


@pytest.mark.parametrize("lat", [D2Q9, D3Q19, D3Q27])
@pytest.mark.parametrize("seed", [3, 4, 5])
def test_newton_matches_bisection(oracle, lat, seed):
    D = 2 if lat == D2Q9 else 3
    f, fneq = cell_state(oracle, lat, 1.01, np.full(D, 0.02), 0.05, seed)
    w = oracle.weights(lat)
    a, b, it = oracle.elbm_alpha(lat, f, fneq, 2.0)
    assert b == oracle.ALPHA_NEWTON and 1 <= it <= 50
    amax = min([2.5] + [abs(f[i]) / fneq[i] for i in range(f.size) if fneq[i] > 0])
    root = bisect_root(lat, w, f, fneq, 1.0 + 1e-9, amax - 1e-12)
    assert abs(a - root) < 2e-8
    # the entropic step stays near the BGK value 2 for small deviations
    assert 1.5 < a < 2.5
    # G(0) = 0 exactly and G changes sign across the root
    assert oracle.entropic_g(lat, f, fneq, 0.0) == 0.0
    dg = oracle.entropic_dg(lat, f, fneq, root)
    assert abs(oracle.entropic_g(lat, f, fneq, root)) < 1e-12 * max(1.0, abs(dg))


def test_newton_returns_iterate_before_last_step(oracle):
    """EntropicStep.h:119-131: alpha is the iterate at which |step| <= tol was
    observed, not alpha - step.  Restate the loop and compare bitwise."""
    lat = D2Q9
    f, fneq = cell_state(oracle, lat, 1.0, np.array([0.02, -0.01]), 0.05, 6)
    x, step = 2.0, 0.0
    for _ in range(50):
        x = x - step
        g = oracle.entropic_g(lat, f, fneq, x)
        step = g / oracle.entropic_dg(lat, f, fneq, x)
        if abs(step) <= 1e-8:
            break
    a, b, _ = oracle.elbm_alpha(lat, f, fneq, 2.0)
    assert b == oracle.ALPHA_NEWTON and a == x


def test_fallback_branch(oracle):
    """A guess near the trivial root alpha = 0 converges there; 0 is outside
    (1, alphaMax) so the reference falls back to alpha = 2 (Collision.h:344-346)."""
    lat = D2Q9
    f, fneq = cell_state(oracle, lat, 1.0, np.array([0.0, 0.01]), 0.05, 7)
    a, b, it = oracle.elbm_alpha(lat, f, fneq, 0.01)
    assert (a, b) == (2.0, oracle.ALPHA_FALLBACK) and it >= 1


@pytest.mark.parametrize("lat", [D2Q9, D3Q27])
def test_elbm_equals_bgk_when_alpha_is_two(oracle, lat):
    nz = 1 if lat == D2Q9 else 4
    kw = dict(lattice=lat, nx=6, ny=5, nz=nz, tau=0.6, force=oracle.FORCE_CONSTANT,
              amplitude=(1e-6, 0, 0))
    cb = oracle.Config(collision=oracle.BGK, **kw)
    ce = oracle.Config(collision=oracle.ELBM, **kw)
    u = oracle.noise_velocity(cb, 11, 1e-5)
    f0 = oracle.init_equilibrium(cb, np.ones(cb.cells), u)
    rb = oracle.run(cb, f0, 4)
    re = oracle.run(ce, f0, 4)
    assert re.branch[oracle.ALPHA_SMALL] == 4 * ce.cells
    np.testing.assert_allclose(re.f, rb.f, rtol=1e-14)


def test_elbm_run_statistics(oracle):
    cfg = oracle.Config(lattice=D2Q9, collision=oracle.ELBM, nx=16, ny=16, tau=0.5003,
                        force=oracle.FORCE_KOLMOGOROV, amplitude=(1e-5, 0, 0),
                        wavelength=(4.0, 1, 1))
    u = oracle.noise_velocity(cfg, 42, 0.03)
    f0 = oracle.init_equilibrium(cfg, np.ones(cfg.cells), u)
    r = oracle.run(cfg, f0, 5)
    assert r.branch.sum() == 5 * cfg.cells
    assert r.branch[oracle.ALPHA_NEWTON] > 0
    assert abs(r.f.sum() - f0.sum()) < 1e-12 * f0.sum()
    assert np.all((r.alpha > 0) & (r.alpha <= 2.5))
