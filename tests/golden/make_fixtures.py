"""Writes the golden fixtures of the fused step from the FP64 oracle.

    python tests/golden/make_fixtures.py

The reference itself cannot be built here (fftw3-mpi.h and the undefined
configuration symbols of its inputs), so these fixtures are the oracle's own
outputs, committed so that (1) the oracle is pinned against regressions of
its restatement and (2) the GPU library can be checked against fixed vectors
without running the oracle.  Every case starts from f = feq(1, u0) with u0
the SplitMix64 noise field of `seed` and amplitude `amp` (oracle
noise_velocity / init_equilibrium), then runs `steps` steps.

Outputs (next to this script):
  steps.npz     per case: final f [Q][N], rho, u_h, alpha, ELBM branch counts
  c1_sha256.txt SHA-256 of the final f (little-endian float64, [Q][N]) of the
                D2Q9 BGK 128^2 Guo + Kolmogorov, 1000-step parity configuration
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as o  # noqa: E402  (test infrastructure)

# name: (lattice, collision, forcing scheme, force, equilibrium, shape, tau,
#        amplitude, wavelength, steps, seed, amp)
D2, D19, D27 = o.D2Q9, o.D3Q19, o.D3Q27
BGK, ELBM = o.BGK, o.ELBM
NONE, GUO, SC, EDM = o.FS_NONE, o.FS_GUO, o.FS_SHANCHEN, o.FS_EDM
F_NONE, F_CONST, F_SIN, F_KOL = (o.FORCE_NONE, o.FORCE_CONSTANT, o.FORCE_SINUSOIDAL,
                                 o.FORCE_KOLMOGOROV)
MA3, EXACT = o.EQ_MA3, o.EQ_EXACT
CASES = {
    "d2q9-bgk-guo-kolmogorov": (D2, BGK, GUO, F_KOL, MA3, (8, 6, 1), 0.6, (1e-3, 0, 0), (6.0, 1, 1), 20, 42, 2e-2),
    "d2q9-bgk-edm-sinusoidal-exact": (D2, BGK, EDM, F_SIN, EXACT, (8, 6, 1), 0.7, (1e-3, -2e-3, 0), (8.0, 6.0, 1), 20, 7, 2e-2),
    "d2q9-bgk-shanchen-constant": (D2, BGK, SC, F_CONST, MA3, (8, 6, 1), 0.8, (1e-4, 2e-4, 0), (1, 1, 1), 20, 3, 2e-2),
    "d3q19-bgk-guo-constant": (D19, BGK, GUO, F_CONST, MA3, (4, 5, 6), 0.55, (1e-4, 0, 2e-4), (1, 1, 1), 10, 42, 2e-2),
    "d3q19-bgk-edm-kolmogorov": (D19, BGK, EDM, F_KOL, MA3, (4, 5, 6), 0.9, (2e-4, 0, 0), (5.0, 1, 1), 10, 11, 2e-2),
    "d3q19-bgk-none": (D19, BGK, NONE, F_NONE, MA3, (4, 5, 6), 0.6, (0, 0, 0), (1, 1, 1), 10, 5, 2e-2),
    "d3q27-bgk-guo-sinusoidal-exact": (D27, BGK, GUO, F_SIN, EXACT, (4, 4, 5), 0.6, (1e-3, 1e-3, -1e-3), (4.0, 4.0, 5.0), 10, 42, 2e-2),
    "d3q27-bgk-shanchen-kolmogorov": (D27, BGK, SC, F_KOL, MA3, (4, 4, 5), 0.51, (1e-3, 0, 0), (4.0, 1, 1), 10, 9, 2e-2),
    "d2q9-elbm-guo-kolmogorov": (D2, ELBM, GUO, F_KOL, MA3, (8, 6, 1), 0.5003, (1e-5, 0, 0), (6.0, 1, 1), 20, 7, 3e-2),
    "d3q19-elbm-none": (D19, ELBM, NONE, F_NONE, MA3, (4, 5, 6), 0.5003, (0, 0, 0), (1, 1, 1), 10, 7, 3e-2),
    "d3q27-elbm-edm-constant": (D27, ELBM, EDM, F_CONST, MA3, (4, 4, 5), 0.5003, (1e-5, 0, 0), (1, 1, 1), 10, 7, 3e-2),
}
C1 = (D2, BGK, GUO, F_KOL, MA3, (128, 128, 1), 0.6, (1e-5, 0, 0), (32.0, 1, 1), 1000, 42, 1e-3)


def config(case):
    lat, col, fs, force, eq, shape, tau, amp, wl, _, _, _ = case
    return o.Config(lattice=lat, collision=col, forcing_scheme=fs, force=force, equilibrium=eq,
                    nx=shape[0], ny=shape[1], nz=shape[2], tau=tau, amplitude=amp,
                    wavelength=wl)


def run_case(case, threads=1):
    cfg = config(case)
    steps, seed, amp = case[9], case[10], case[11]
    u0 = o.noise_velocity(cfg, seed, amp)
    f0 = o.init_equilibrium(cfg, np.ones(cfg.cells), u0)
    return o.run(cfg, f0, steps, threads=threads)


def sha256_of(f: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(f, dtype="<f8").tobytes()).hexdigest()


def main() -> int:
    o.build()
    out = {}
    for name, case in CASES.items():
        r = run_case(case)
        out[f"{name}/f"] = r.f
        out[f"{name}/rho"] = r.rho
        out[f"{name}/u"] = r.u
        out[f"{name}/alpha"] = r.alpha
        out[f"{name}/branch"] = r.branch
        print(f"{name}: f sum {r.f.sum():.15e}, branches {r.branch.tolist()}")
    np.savez_compressed(os.path.join(HERE, "steps.npz"), **out)
    r = run_case(C1, threads=min(8, os.cpu_count() or 1))
    with open(os.path.join(HERE, "c1_sha256.txt"), "w") as fh:
        fh.write(sha256_of(r.f) + "\n")
    print("c1", sha256_of(r.f))
    with open(os.path.join(HERE, "cases.json"), "w") as fh:
        json.dump({k: list(v[:5]) + [list(v[5]), v[6], list(v[7]), list(v[8])] + list(v[9:])
                   for k, v in {**CASES, "c1": C1}.items()}, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
