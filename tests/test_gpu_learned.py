"""Learned entropic alpha on the GPU, checked against the FP64 oracle.

* The training kernels' gradient and the inference code's prediction match
  the FP64 reference network (oracle/mlp_ref.c) on the same weights.
* End to end: (features, alpha) pairs are logged from the GPU's own Newton
  solves, the network is trained on the GPU, and learned-alpha steps are
  checked cell by cell: every Newton-branch cell either satisfies the
  H-residual bound |G(alpha)| <= tol with G evaluated by the oracle
  (EntropicStep.h:31-45 restated in cpu_ref.c) on the same pulled state, or
  carries the oracle's reference-Newton alpha from the same stored alpha
  (the fallback).
* Ablation: the same certified correction started from the stored alpha
  (the method without the network) on the same state; the test reports what
  the network changes in evaluations per solved cell.
"""
import sys

import numpy as np
import pytest

from metalbm_amd import (AlphaMode, Algorithm, CollisionType, Config, ForceType,
                         ForcingSchemeType, LatticeType, Precision)

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("lat", [LatticeType.D2Q9, LatticeType.D3Q27])
def test_mlp_gradient_and_forward_match_oracle(oracle, lat):
    shape = (8, 8, 1) if lat == LatticeType.D2Q9 else (4, 4, 4)
    cfg = Config(lattice=lat, collision=CollisionType.ELBM, nx=shape[0], ny=shape[1],
                 nz=shape[2], tau=0.6)
    say(f"case mlp grad {lat.name}")
    with Algorithm(cfg) as a:
        a.mlp_init(7)
        nin, h1, h2, npar = a.mlp_shape()
        rng = np.random.default_rng(3)
        mean = (0.1 * rng.standard_normal(nin)).astype(np.float32)
        scale = rng.uniform(0.5, 2.0, nin).astype(np.float32)
        a.set_normalization(mean, scale, 2.0, 1e-3)
        p = a.mlp_params()
        # perturb biases so both ReLU sides are exercised
        p = (p + 0.05 * rng.standard_normal(npar)).astype(np.float32)
        a.set_mlp_params(p)
        x = rng.standard_normal((96, nin)).astype(np.float32)
        alpha = 2.0 + 1e-3 * rng.standard_normal(96)
        g, loss = a.mlp_gradient(x, alpha)
        pred = a.mlp_predict(x)
    z = (x.astype(np.float64) - mean.astype(np.float64)) * scale.astype(np.float64)
    t = (alpha - 2.0) / np.float64(np.float32(1e-3))
    lref, gref = oracle.mlp_grad(p.astype(np.float64), nin, h1, h2, z, t)
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    err = np.abs(g.astype(np.float64) - gref)
    assert np.all(err <= 1e-4 * np.abs(gref).max() + 1e-4 * np.abs(gref)), err.max()
    yref = oracle.mlp_forward(p.astype(np.float64), nin, h1, h2, z)
    # Inference runs on the matrix cores with f16 (D2Q9) or bf16 (D3Q27)
    # operands (features, weights, hidden activations) and FP32 accumulation;
    # biases stay FP32.  Forward rounding-error bound with unit roundoff u per
    # operand rounding (f16: 2^-11, bf16: 2^-8), layer by layer:
    #   e1 = 2u |W1||z| + u |a1|,  e2 = |W2| e1 + 2u |W2||a1| + u |a2|,
    #   e3 = |W3| e2 + 2u |W3||a2|   (ReLU is 1-Lipschitz), doubled for slack.
    u = 2.0 ** -11 if lat == LatticeType.D2Q9 else 2.0 ** -8
    P = p.astype(np.float64)
    W1 = np.abs(P[:h1 * nin].reshape(h1, nin))
    o = h1 * nin + h1
    W2 = np.abs(P[o:o + h2 * h1].reshape(h2, h1))
    W3 = np.abs(P[o + h2 * h1 + h2:o + h2 * h1 + 2 * h2])
    b1 = P[h1 * nin:h1 * nin + h1]
    b2 = P[o + h2 * h1:o + h2 * h1 + h2]
    a1 = np.maximum(z @ P[:h1 * nin].reshape(h1, nin).T + b1, 0)
    a2 = np.maximum(a1 @ P[o:o + h2 * h1].reshape(h2, h1).T + b2, 0)
    e1 = 2 * u * (np.abs(z) @ W1.T) + u * a1
    e2 = e1 @ W2.T + 2 * u * (a1 @ W2.T) + u * a2
    e3 = e2 @ W3 + 2 * u * (a2 @ W3)
    out_scale = np.float64(np.float32(1e-3))
    bound = 2.0 * out_scale * e3 + 1e-12
    err = np.abs(pred - (2.0 + out_scale * yref))
    say(f"matrix-core inference |err| max {err.max():.3e} median {np.median(err):.3e}; "
        f"bound median {np.median(bound):.3e}; output range {out_scale * np.ptp(yref):.3e}")
    assert np.all(err <= bound), (err / bound).max()


def pulled_state(o, oc, f):
    """Pulled populations and fneq = f - feq(rho, u) per cell (oracle order)."""
    c = o.velocities(oc.lattice)
    Q, N = f.shape
    shp = (oc.nx, oc.ny, oc.nz)
    pad = (0,) if oc.d == 2 else ()
    fp = np.stack([np.roll(f[i].reshape(shp), tuple(int(v) for v in c[i]) + pad,
                           axis=(0, 1, 2)).ravel() for i in range(Q)])
    rho = fp.sum(axis=0)
    u = (c.T.astype(float) @ fp) / rho
    feq = np.stack([o.equilibrium(oc, rho[k], u[:, k]) for k in range(N)], axis=1)
    return fp, fp - feq


CASES = [(LatticeType.D2Q9, (64, 64, 1), 1e-3, 16.0),
         (LatticeType.D3Q27, (12, 12, 12), 3e-3, 6.0)]


@pytest.mark.parametrize("lat,shape,amp,lam", CASES, ids=["D2Q9-64x64", "D3Q27-12^3"])
def test_learned_alpha_end_to_end(oracle, lat, shape, amp, lam):
    nx, ny, nz = shape
    tol = 1e-10
    cfg = Config(lattice=lat, collision=CollisionType.ELBM, forcing_scheme=ForcingSchemeType.Guo,
                 force=ForceType.Kolmogorov, nx=nx, ny=ny, nz=nz, tau=0.50003,
                 amplitude=(1e-5, 0, 0), wavelength=(lam, 1, 1), residual_tol=tol)
    oc = oracle.Config(lattice=int(lat), collision=2, forcing_scheme=2, force=6, equilibrium=3,
                       nx=nx, ny=ny, nz=nz, tau=0.50003, amplitude=(1e-5, 0, 0),
                       wavelength=(lam, 1, 1))
    say(f"case learned e2e {lat.name} {shape}")
    with Algorithm(cfg) as a:
        a.init_noise(1.0, amp, 42)
        a.sample_log_enable(200000, 1.0, 5)
        a.iterate(60)
        a.sample_log_disable()
        x, al = a.sample_log()
        say(f"logged {len(al)} samples, alpha-2 std {np.std(al - 2):.3e}")
        assert len(al) >= 1000
        l0, l1 = a.mlp_train(1500, 1024, 2e-3, seed=1)
        say(f"training loss {l0:.4e} -> {l1:.4e}")
        assert l1 < 0.25 * l0
        # training is reproducible: same samples and seed, same network
        p1 = a.mlp_params()
        a.mlp_init(1)
        assert a.mlp_train(1500, 1024, 2e-3, seed=1) == (l0, l1)
        np.testing.assert_array_equal(a.mlp_params(), p1)
        pred = a.mlp_predict(x[:20000])
        err = np.abs(pred - al[:20000])
        spread = np.std(al - 2)
        say(f"prediction |err| median {np.median(err):.3e} p95 {np.percentile(err, 95):.3e} "
            f"vs spread {spread:.3e}")
        assert np.median(err) < 0.25 * spread
        f0 = a.populations()
        alpha0 = a.alpha()
        res = {}
        for mode in (AlphaMode.NEWTON, AlphaMode.WARM_RESIDUAL, AlphaMode.LEARNED):
            a.set_populations(f0)
            a.set_alpha(alpha0)
            a.reset_elbm_stats()
            a.set_alpha_mode(mode, tol)
            a.iterate(1, store=True)
            res[mode] = (a.store_fields()[2], a.elbm_stats(), a.populations())
    fp, fneq = pulled_state(oracle, oc, f0)
    newton_cells, ref_fallbacks, worst, by_ref = 0, 0, 0.0, 0
    for k in range(fp.shape[1]):
        aref, b, _ = oracle.elbm_alpha(int(lat), fp[:, k], fneq[:, k], alpha0[k])
        if b in (oracle.ALPHA_SMALL, oracle.ALPHA_CAPPED):
            continue
        newton_cells += 1
        ref_fallbacks += b != oracle.ALPHA_NEWTON
        for mode in (AlphaMode.WARM_RESIDUAL, AlphaMode.LEARNED):
            al = res[mode][0][k]
            g = abs(oracle.entropic_g(int(lat), fp[:, k], fneq[:, k], al))
            if g <= 1.001 * tol:
                worst = max(worst, g)
            else:  # the reference Newton decided this cell
                assert abs(al - aref) <= 1e-9 * abs(aref), (k, mode, al, aref, g)
                by_ref += 1
    say(f"{newton_cells} solved cells ({ref_fallbacks} oracle fallbacks); worst accepted "
        f"|G(alpha)| {worst:.3e} (tol {tol:.0e}); {by_ref} cell-solves left to the reference")
    assert newton_cells > 0.5 * fp.shape[1]
    for mode, (_, st, _) in res.items():
        per = st["newton_iterations"] / max(1, st["newton"])
        say(f"mode {mode.name}: {st}; evaluations per solved cell {per:.3f}")
    st_l, st_w = res[AlphaMode.LEARNED][1], res[AlphaMode.WARM_RESIDUAL][1]
    assert st_l["fallback"] <= ref_fallbacks
    # all three solvers find the same root to within the residual tolerance
    fN, fL = res[AlphaMode.NEWTON][2], res[AlphaMode.LEARNED][2]
    assert np.max(np.abs(fL - fN)) <= 1e-6 * np.max(np.abs(fN))
    # what the network changes: evaluations per solved cell vs the same solver
    # started from the stored alpha
    ev_l = st_l["newton_iterations"] / max(1, st_l["newton"])
    ev_w = st_w["newton_iterations"] / max(1, st_w["newton"])
    say(f"network gain: {ev_w:.3f} -> {ev_l:.3f} evaluations per solved cell")
    assert ev_l <= ev_w + 1e-9
