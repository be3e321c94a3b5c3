"""Spectral forces on the GPU (ConstantShell, EnergyRemoval, Turbulent2D,
Force.h:294-623): the force field against the numpy oracle, one step against
the C oracle fed the same field (bitwise for FP64 BGK), and whole runs with
the reference's update order (field recomputed before every step from the
fields of the last storing step) against the oracle's own fields.  The two
FFT implementations round differently, so field and run comparisons are
relative to the largest magnitude."""
import numpy as np
import pytest

from metalbm_amd import (Algorithm, CollisionType, Config, ForceType, ForcingSchemeType,
                         LatticeType, Precision)
from oracle import oracle as o

pytestmark = pytest.mark.gpu

SPEC = dict(force_kmin=2, force_kmax=4, removal_amplitude=(1e-2, 2e-2, 1e-2),
            removal_kmin=1, removal_kmax=3)
CASES = [
    ("d2q9-turbulent2d-edm", Config(lattice=LatticeType.D2Q9, nx=32, ny=24,
                                    forcing_scheme=ForcingSchemeType.ExactDifferenceMethod,
                                    force=ForceType.Turbulent2D, amplitude=(5e-3, 0, 0),
                                    tau=0.6, **SPEC)),
    ("d2q9-elbm-turbulent2d", Config(lattice=LatticeType.D2Q9, nx=24, ny=20,
                                     collision=CollisionType.ELBM,
                                     forcing_scheme=ForcingSchemeType.ExactDifferenceMethod,
                                     force=ForceType.Turbulent2D, amplitude=(5e-3, 0, 0),
                                     tau=0.5003, **SPEC)),
    ("d3q19-constant-shell-guo", Config(lattice=LatticeType.D3Q19, nx=16, ny=12, nz=10,
                                        force=ForceType.ConstantShell,
                                        amplitude=(1e-2, 2e-2, -1e-2), tau=0.6, **SPEC)),
    ("d3q27-energy-removal", Config(lattice=LatticeType.D3Q27, nx=12, ny=10, nz=8,
                                    force=ForceType.EnergyRemoval, tau=0.7, **SPEC)),
    ("d3q27-turbulent-odd", Config(lattice=LatticeType.D3Q27, nx=9, ny=10, nz=7,
                                   forcing_scheme=ForcingSchemeType.ExactDifferenceMethod,
                                   force=ForceType.Turbulent2D, amplitude=(1e-2, 0, 1e-2),
                                   tau=0.8, **SPEC)),
    # removal shells of <= 32 modes take the direct transform (2D above);
    # these cover the direct path in 3D, the 2D path for |k_d| <= 2 and the
    # full transforms in 2D
    ("d3q19-energy-removal-direct", Config(lattice=LatticeType.D3Q19, nx=10, ny=12, nz=8,
                                           force=ForceType.EnergyRemoval, tau=0.7,
                                           **{**SPEC, "removal_kmax": 2})),
    # 2D shell with every |k_d| <= 2: the dense separable synthesis
    ("d2q9-turbulent2d-direct-small", Config(lattice=LatticeType.D2Q9, nx=32, ny=24,
                                             forcing_scheme=ForcingSchemeType.ExactDifferenceMethod,
                                             force=ForceType.Turbulent2D, amplitude=(5e-3, 0, 0),
                                             tau=0.6, **{**SPEC, "removal_kmax": 2})),
    ("d2q9-energy-removal-direct-small", Config(lattice=LatticeType.D2Q9, nx=20, ny=18,
                                                force=ForceType.EnergyRemoval, tau=0.7,
                                                **{**SPEC, "removal_kmax": 2})),
    ("d2q9-turbulent2d-fft", Config(lattice=LatticeType.D2Q9, nx=30, ny=28,
                                    forcing_scheme=ForcingSchemeType.ExactDifferenceMethod,
                                    force=ForceType.Turbulent2D, amplitude=(5e-3, 0, 0),
                                    tau=0.6, **{**SPEC, "removal_kmax": 6})),
]
IDS = [c[0] for c in CASES]


def oracle_cfg(c):
    return o.Config(lattice=int(c.lattice), collision=int(c.collision),
                    forcing_scheme=int(c.forcing_scheme), force=int(c.force),
                    equilibrium=int(c.equilibrium), nx=c.nx, ny=c.ny, nz=c.nz, tau=c.tau,
                    amplitude=tuple(c.amplitude), wavelength=tuple(c.wavelength))


def spec_of(c):
    return dict(kmin=c.force_kmin, kmax=c.force_kmax, removal_amplitude=c.removal_amplitude,
                removal_kmin=c.removal_kmin, removal_kmax=c.removal_kmax)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name,c", CASES, ids=IDS)
def test_force_field_matches_oracle(name, c):
    print(f"case spectral force field {name}", flush=True)
    oc = oracle_cfg(c)
    u0 = o.noise_velocity(oc, 7, 2e-2)
    with Algorithm(c) as a:
        a.init_noise(1.0, 2e-2, 7)
        a.iterate(1, store=True)
        F1 = a.store_fields()[3]
        Fo = o.spectral_force_field(oc, spec_of(c), np.ones(oc.cells), u0)
        assert np.max(np.abs(Fo)) > 0
        assert rel(F1, Fo) <= 1e-12, rel(F1, Fo)
        # the next field comes from the fields of that storing step
        rho, u = a.store_fields()[:2]
        a.iterate(1, store=True)
        F2 = a.store_fields()[3]
        Fo2 = o.spectral_force_field(oc, spec_of(c), rho, u)
        assert rel(F2, Fo2) <= 1e-12, rel(F2, Fo2)


@pytest.mark.parametrize("name,c", CASES, ids=IDS)
def test_step_with_the_gpu_field_matches_oracle(name, c):
    print(f"case spectral force step {name}", flush=True)
    oc = oracle_cfg(c)
    with Algorithm(c) as a:
        a.init_noise(1.0, 2e-2, 3)
        for _ in range(4):
            a.iterate(1, store=True)
        f4 = a.populations()
        al4 = a.alpha() if c.collision == CollisionType.ELBM else None
        a.iterate(1, store=True)
        f5 = a.populations()
        F5 = a.store_fields()[3]
    fo, rho, u, Fo = o.step_field(oc, f4, al4, F5)
    np.testing.assert_array_equal(Fo, F5)
    if c.collision == CollisionType.BGK:
        np.testing.assert_array_equal(f5, fo)  # FP64 reference order: bitwise
    else:
        assert np.max(np.abs(f5 - fo)) <= 1e-12


@pytest.mark.parametrize("store_every", [1, 5])
@pytest.mark.parametrize("name,c", CASES, ids=IDS)
def test_runs_match_oracle(name, c, store_every):
    print(f"case spectral force run {name} store_every={store_every}", flush=True)
    oc = oracle_cfg(c)
    u0 = o.noise_velocity(oc, 11, 2e-2)
    f0 = o.init_equilibrium(oc, np.ones(oc.cells), u0)
    fo, alo, rho_o, u_o, F_o = o.run_field_force(oc, f0, 10, np.ones(oc.cells), u0, spec_of(c),
                                                 store_every=store_every)
    with Algorithm(c) as a:
        a.init_noise(1.0, 2e-2, 11)
        for _ in range(10 // store_every):
            a.iterate(store_every, store=True)
        f = a.populations()
        rho, u, al, F = a.store_fields()
    assert rel(F, F_o) <= 1e-11, rel(F, F_o)
    # BGK differs only through the force fields' rounding; ELBM adds the
    # device log's (the entropic parity bar is 1e-12)
    tol = 1e-12 if c.collision == CollisionType.ELBM else 1e-13
    assert np.max(np.abs(f - fo)) <= tol, np.max(np.abs(f - fo))
    assert rel(u, u_o) <= 1e-10, rel(u, u_o)


def test_fp32_compute_tracks_fp64():
    print("case spectral force fp32", flush=True)
    c64 = CASES[0][1]
    c32 = Config(**{**c64.__dict__, "storage": Precision.F32, "compute": Precision.F32,
                    "shifted": True})
    out = []
    for c in (c64, c32):
        with Algorithm(c) as a:
            a.init_noise(1.0, 2e-2, 5)
            for _ in range(4):
                a.iterate(5, store=True)
            out.append(a.store_fields())
    assert rel(out[1][3], out[0][3]) <= 1e-5   # force field
    assert rel(out[1][1], out[0][1]) <= 1e-4   # velocity


@pytest.mark.parametrize("name,c", [CASES[0], CASES[1], CASES[3]],
                         ids=[IDS[0], IDS[1], IDS[3]])
def test_reinitialised_handle_matches_a_fresh_one(name, c):
    """The force field is rebuilt only when the fields it reads change (a
    storing step, an initialiser, a restart), and an initialiser restarts the
    entropic alpha field at 2 (initAlpha, Initialize.h:81-88).  A handle
    re-initialised in the middle of a run must give the same steps, bit for
    bit, as a fresh handle given the same initial field."""
    print(f"case spectral force re-init {name}", flush=True)
    with Algorithm(c) as a:
        a.init_noise(1.0, 2e-2, 3)
        a.iterate(4, store=True)
        a.iterate(3)                      # non-storing steps reuse the field
        a.init_noise(1.0, 2e-2, 9)
        a.iterate(6, store=True)
        f_re = a.populations()
        F_re = a.store_fields()[3]
    with Algorithm(c) as b:
        b.init_noise(1.0, 2e-2, 9)
        b.iterate(6, store=True)
        f_fresh = b.populations()
        F_fresh = b.store_fields()[3]
    np.testing.assert_array_equal(F_re, F_fresh)
    np.testing.assert_array_equal(f_re, f_fresh)


def test_restart_from_file_resets_alpha(tmp_path):
    """A restart from a distribution file starts the entropic solve from
    alpha = 2, as the reference's initAlpha does on every start
    (Initialize.h:81-88, 119-145)."""
    print("case elbm restart alpha", flush=True)
    c = CASES[1][1]
    with Algorithm(c) as a:
        a.init_noise(1.0, 2e-2, 3)
        a.iterate(5)
        assert np.any(a.alpha() != 2.0)
        a.write_distribution_hdf5(str(tmp_path / "d.h5"))
        a.read_distribution_hdf5(str(tmp_path / "d.h5"))
        np.testing.assert_array_equal(a.alpha(), np.full(c.cells, 2.0))


def test_restart_fields_feed_energy_removal():
    """After lbm_set_populations (a restart), EnergyRemoval reads the fields
    the reference's restart derives from the populations (Initialize.h:
    124-145): density sum_i f_i and velocity sum_i c_i f_i, undivided."""
    print("case spectral force restart fields", flush=True)
    c = CASES[0][1]
    oc = oracle_cfg(c)
    with Algorithm(c) as a:
        a.init_noise(1.0, 2e-2, 13)
        a.iterate(5, store=True)
        f5 = a.populations()
    with Algorithm(c) as b:
        b.set_populations(f5)
        b.iterate(3, store=True)  # all three steps use the restart fields
        f8 = b.populations()
    cv = o.velocities(oc.lattice).astype(np.float64)  # [Q, D]
    rho0 = f5.sum(axis=0)
    mom0 = cv.T @ f5
    fo = o.run_field_force(oc, f5, 3, rho0, mom0, spec_of(c), store_every=3)[0]
    assert np.max(np.abs(f8 - fo)) <= 1e-13, np.max(np.abs(f8 - fo))
    # dividing by the density instead would be visible at this tolerance
    fd = o.run_field_force(oc, f5, 3, rho0, mom0 / rho0, spec_of(c), store_every=3)[0]
    assert np.max(np.abs(fd - fo)) > 1e-12
