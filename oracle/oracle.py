"""ctypes front end of the FP64 CPU oracle (oracle/cpu_ref.c, oracle/mlp_ref.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or the timed CPU baseline); the product path never touches it.

Array conventions follow the C side: populations ``f[Q, N]``, velocities
``u[D, N]`` (component-major), cells ``(x * ny + y) * nz + z``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcpu_ref.so")

# Reference enum values (include/metaLBM/Options.h:12-41).
D2Q9, D3Q19, D3Q27 = 3, 9, 10
BGK, ELBM = 1, 2
FS_NONE, FS_GUO, FS_SHANCHEN, FS_EDM = 1, 2, 3, 4
FORCE_NONE, FORCE_CONSTANT, FORCE_SINUSOIDAL, FORCE_KOLMOGOROV = 0, 4, 5, 6
FORCE_CONSTANT_SHELL, FORCE_ENERGY_REMOVAL, FORCE_TURBULENT_2D = 7, 8, 9
EQ_EXACT, EQ_MA3 = 1, 3
ALPHA_SMALL, ALPHA_CAPPED, ALPHA_NEWTON, ALPHA_FALLBACK = 0, 1, 2, 3


class _Config(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int), ("collision", ctypes.c_int),
        ("forcing_scheme", ctypes.c_int), ("force", ctypes.c_int),
        ("equilibrium", ctypes.c_int),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("tau", ctypes.c_double),
        ("amplitude", ctypes.c_double * 3), ("wavelength", ctypes.c_double * 3),
    ]


class _Stats(ctypes.Structure):
    _fields_ = [("branch", ctypes.c_int64 * 4), ("newton_iterations", ctypes.c_int64)]


def build() -> str:
    """Compile the oracle with its Makefile (gcc, seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None
_lib_native = None
_NATIVE_PATH = os.path.join(_HERE, "_native", "libcpu_ref.so")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def lib_native():
    """The same oracle compiled -march=native on this host (make native):
    the timed CPU baseline, built where it is timed."""
    global _lib_native
    if _lib_native is None:
        subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True)
        _lib_native = _bind(ctypes.CDLL(_NATIVE_PATH))
    return _lib_native


def _bind(L):
    dp = ctypes.POINTER(ctypes.c_double)
    cp = ctypes.POINTER(_Config)
    L.ref_lattice_q.argtypes = [ctypes.c_int]
    L.ref_lattice_d.argtypes = [ctypes.c_int]
    L.ref_lattice_c.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.ref_lattice_w.argtypes = [ctypes.c_int, ctypes.c_int]
    L.ref_lattice_w.restype = ctypes.c_double
    L.ref_equilibrium.argtypes = [cp, ctypes.c_double, dp, ctypes.c_int]
    L.ref_equilibrium.restype = ctypes.c_double
    L.ref_force_at.argtypes = [cp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
    L.ref_init_equilibrium.argtypes = [cp, dp, dp, dp]
    L.ref_step.argtypes = [cp, dp, dp, dp, dp, dp, dp, ctypes.POINTER(_Stats)]
    L.ref_step_field.argtypes = [cp, dp, dp, dp, dp, dp, dp, dp, ctypes.POINTER(_Stats)]
    L.ref_run.argtypes = [cp, dp, dp, dp, ctypes.c_int, dp, dp, dp,
                          ctypes.POINTER(_Stats)]
    L.ref_elbm_alpha.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double,
                                 ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(ctypes.c_int)]
    L.ref_elbm_alpha.restype = ctypes.c_double
    L.ref_entropic_g.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double]
    L.ref_entropic_g.restype = ctypes.c_double
    L.ref_entropic_dg.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double]
    L.ref_entropic_dg.restype = ctypes.c_double
    L.ref_splitmix64.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
    L.ref_splitmix64.restype = ctypes.c_uint64
    L.ref_init_noise_velocity.argtypes = [cp, ctypes.c_uint64, ctypes.c_double, dp]
    L.ref_init_taylor_green.argtypes = [cp, ctypes.c_double, dp]
    L.ref_init_spectral_velocity.argtypes = [cp, ctypes.c_double, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_uint64, dp]
    L.ref_init_spectral_velocity.restype = ctypes.c_int
    L.ref_reduce.argtypes = [cp, dp, dp, dp, dp, dp]
    L.ref_set_threads.argtypes = [ctypes.c_int]
    L.mlp_ref_param_count.argtypes = [ctypes.c_int] * 3
    L.mlp_ref_forward.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, dp, dp]
    L.mlp_ref_loss.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, dp, dp]
    L.mlp_ref_loss.restype = ctypes.c_double
    L.mlp_ref_grad.argtypes = [dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, dp, dp, dp]
    L.mlp_ref_grad.restype = ctypes.c_double
    L.mlp_ref_adam.argtypes = [dp, dp, dp, dp, ctypes.c_int, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.c_int]
    return L


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Config:
    lattice: int = D2Q9
    collision: int = BGK
    forcing_scheme: int = FS_GUO
    force: int = FORCE_NONE
    equilibrium: int = EQ_MA3
    nx: int = 16
    ny: int = 16
    nz: int = 1
    tau: float = 0.6
    amplitude: tuple = (0.0, 0.0, 0.0)
    wavelength: tuple = (1.0, 1.0, 1.0)

    def c(self) -> _Config:
        k = _Config()
        for n in ("lattice", "collision", "forcing_scheme", "force", "equilibrium",
                  "nx", "ny", "nz"):
            setattr(k, n, int(getattr(self, n)))
        k.tau = float(self.tau)
        for d in range(3):
            k.amplitude[d] = float(self.amplitude[d])
            k.wavelength[d] = float(self.wavelength[d])
        return k

    @property
    def q(self) -> int:
        return lib().ref_lattice_q(self.lattice)

    @property
    def d(self) -> int:
        return lib().ref_lattice_d(self.lattice)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz


def velocities(lattice: int) -> np.ndarray:
    L = lib()
    q, d = L.ref_lattice_q(lattice), L.ref_lattice_d(lattice)
    return np.array([[L.ref_lattice_c(lattice, i, k) for k in range(d)]
                     for i in range(q)], dtype=np.int64)


def weights(lattice: int) -> np.ndarray:
    L = lib()
    return np.array([L.ref_lattice_w(lattice, i)
                     for i in range(L.ref_lattice_q(lattice))])


def equilibrium(cfg: Config, rho: float, u) -> np.ndarray:
    u = np.ascontiguousarray(np.asarray(u, dtype=np.float64).ravel())
    k = cfg.c()
    return np.array([lib().ref_equilibrium(ctypes.byref(k), rho, _p(u), i)
                     for i in range(cfg.q)])


def force_at(cfg: Config, x: int, y: int, z: int = 0) -> np.ndarray:
    out = np.zeros(3)
    k = cfg.c()
    lib().ref_force_at(ctypes.byref(k), x, y, z, _p(out))
    return out


def init_equilibrium(cfg: Config, rho: np.ndarray, u: np.ndarray) -> np.ndarray:
    N = cfg.cells
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(N)
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(cfg.d, N)
    f = np.empty((cfg.q, N))
    k = cfg.c()
    lib().ref_init_equilibrium(ctypes.byref(k), _p(rho), _p(u), _p(f))
    return f


def noise_velocity(cfg: Config, seed: int, amp: float) -> np.ndarray:
    u = np.zeros((cfg.d, cfg.cells))
    k = cfg.c()
    lib().ref_init_noise_velocity(ctypes.byref(k), seed, amp, _p(u))
    return u


def taylor_green(cfg: Config, u0: float) -> np.ndarray:
    u = np.zeros((cfg.d, cfg.cells))
    k = cfg.c()
    lib().ref_init_taylor_green(ctypes.byref(k), u0, _p(u))
    return u


def spectral_velocity(cfg: Config, urms: float, kmax: int, kpeak: float,
                      seed: int) -> np.ndarray:
    """Solenoidal random-phase field with E(s) = s^4 exp(-2 (s/kpeak)^2) on
    shells 1..kmax and per-component rms `urms` (cpu_ref.h states the spec)."""
    u = np.zeros((cfg.d, cfg.cells))
    k = cfg.c()
    rc = lib().ref_init_spectral_velocity(ctypes.byref(k), urms, kmax, kpeak, seed, _p(u))
    if rc != 0:
        raise ValueError("spectral init needs 1 <= kmax and 2 kmax < every extent")
    return u


@dataclass
class RunResult:
    f: np.ndarray
    alpha: np.ndarray
    rho: np.ndarray
    u: np.ndarray
    force: np.ndarray
    branch: np.ndarray = field(default_factory=lambda: np.zeros(4, np.int64))
    newton_iterations: int = 0


def run(cfg: Config, f0: np.ndarray, nsteps: int, alpha0: np.ndarray | None = None,
        threads: int = 1, native: bool = False) -> RunResult:
    """Step the oracle nsteps times from f0; fields are those of the last step.
    native: run the -march=native build of the same sources (timing only)."""
    N, Q, D = cfg.cells, cfg.q, cfg.d
    f = np.array(f0, dtype=np.float64, order="C").reshape(Q, N).copy()
    scratch = np.empty_like(f)
    alpha = (np.full(N, 2.0) if alpha0 is None
             else np.array(alpha0, dtype=np.float64).reshape(N).copy())
    rho = np.zeros(N)
    u = np.zeros((D, N))
    F = np.zeros((D, N))
    st = _Stats()
    L = lib_native() if native else lib()
    L.ref_set_threads(threads)
    k = cfg.c()
    rc = L.ref_run(ctypes.byref(k), _p(f), _p(scratch), _p(alpha), nsteps,
                   _p(rho), _p(u), _p(F), ctypes.byref(st))
    if rc != 0:
        raise ValueError(f"oracle rejected the configuration (code {rc})")
    return RunResult(f, alpha, rho, u, F, np.array(st.branch[:], dtype=np.int64),
                     int(st.newton_iterations))


def step_field(cfg: Config, f: np.ndarray, alpha: np.ndarray | None, force_field: np.ndarray,
               threads: int = 1):
    """One step with the force read from force_field[D, N] (the Generic
    setForce, Force.h:39-48).  Returns (f_next, rho, u_hydro, force); alpha
    (ELBM warm start) is updated in place."""
    N, Q, D = cfg.cells, cfg.q, cfg.d
    prev = np.ascontiguousarray(f, dtype=np.float64).reshape(Q, N)
    nxt = np.empty_like(prev)
    ff = np.ascontiguousarray(force_field, dtype=np.float64).reshape(D, N)
    if alpha is None:
        alpha = np.full(N, 2.0)
    rho, u, F = np.zeros(N), np.zeros((D, N)), np.zeros((D, N))
    L = lib()
    L.ref_set_threads(threads)
    k = cfg.c()
    rc = L.ref_step_field(ctypes.byref(k), _p(prev), _p(nxt), _p(alpha), _p(ff), _p(rho),
                          _p(u), _p(F), None)
    if rc != 0:
        raise ValueError(f"oracle rejected the configuration (code {rc})")
    return nxt, rho, u, F


def spectral_force_field(cfg: Config, spec: dict, rho: np.ndarray, u: np.ndarray) -> np.ndarray:
    """The force field of a spectral force (oracle/spectral.py) from the
    stored fields rho[N], u[D, N].  spec: kmin, kmax, removal_amplitude,
    removal_kmin, removal_kmax (the input file's forcekMin ... names)."""
    from oracle import spectral as sp
    shape = (cfg.nx, cfg.ny) if cfg.d == 2 else (cfg.nx, cfg.ny, cfg.nz)
    if cfg.force == FORCE_CONSTANT_SHELL:
        return sp.constant_shell_force(shape, cfg.amplitude, spec["kmin"], spec["kmax"])
    if cfg.force == FORCE_ENERGY_REMOVAL:
        return sp.energy_removal_force(rho, u, shape, spec["removal_amplitude"],
                                       spec["removal_kmin"], spec["removal_kmax"])
    if cfg.force == FORCE_TURBULENT_2D:
        return sp.turbulent_force(rho, u, shape, cfg.amplitude, spec["kmin"], spec["kmax"],
                                  spec["removal_amplitude"], spec["removal_kmin"],
                                  spec["removal_kmax"])
    raise ValueError("not a spectral force")


def run_field_force(cfg: Config, f0: np.ndarray, nsteps: int, rho0: np.ndarray,
                    u0: np.ndarray, spec: dict, store_every: int = 1,
                    alpha0: np.ndarray | None = None, threads: int = 1):
    """Steps with a spectral force as Algorithm::iterate runs them: the force
    field is recomputed before every step (Collision::update, Algorithm.h:338;
    ConstantShell once) from the fields of the last storing step, which are
    the initial fields (rho0, u0) until the first (Initialize.h:20-70); every
    store_every-th step stores.  Returns (f, alpha, rho, u, F) of the last
    step."""
    N, Q, D = cfg.cells, cfg.q, cfg.d
    f = np.array(f0, dtype=np.float64).reshape(Q, N)
    alpha = (np.full(N, 2.0) if alpha0 is None
             else np.array(alpha0, dtype=np.float64).reshape(N).copy())
    rho_s = np.array(rho0, dtype=np.float64).reshape(N)
    u_s = np.array(u0, dtype=np.float64).reshape(D, N)
    F = None
    out = None
    for step in range(1, nsteps + 1):
        if F is None or cfg.force != FORCE_CONSTANT_SHELL:
            F = spectral_force_field(cfg, spec, rho_s, u_s)
        f, rho, u, Fo = step_field(cfg, f, alpha, F, threads)
        if step % store_every == 0:
            rho_s, u_s = rho, u
        out = (rho, u, Fo)
    return f, alpha, out[0], out[1], out[2]


def elbm_alpha(lattice: int, f, fneq, alpha_guess: float = 2.0):
    f = np.ascontiguousarray(f, dtype=np.float64)
    fneq = np.ascontiguousarray(fneq, dtype=np.float64)
    b = ctypes.c_int()
    it = ctypes.c_int()
    a = lib().ref_elbm_alpha(lattice, _p(f), _p(fneq), alpha_guess,
                             ctypes.byref(b), ctypes.byref(it))
    return a, b.value, it.value


def entropic_g(lattice: int, f, fneq, alpha: float) -> float:
    f = np.ascontiguousarray(f, dtype=np.float64)
    fneq = np.ascontiguousarray(fneq, dtype=np.float64)
    return lib().ref_entropic_g(lattice, _p(f), _p(fneq), alpha)


def entropic_dg(lattice: int, f, fneq, alpha: float) -> float:
    f = np.ascontiguousarray(f, dtype=np.float64)
    fneq = np.ascontiguousarray(fneq, dtype=np.float64)
    return lib().ref_entropic_dg(lattice, _p(f), _p(fneq), alpha)


def splitmix64(seed: int, k: int) -> int:
    return int(lib().ref_splitmix64(seed, k))


def reduce(cfg: Config, rho: np.ndarray, u: np.ndarray):
    rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(cfg.cells)
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(cfg.d, cfg.cells)
    m, e, w = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    k = cfg.c()
    lib().ref_reduce(ctypes.byref(k), _p(rho), _p(u), ctypes.byref(m),
                     ctypes.byref(e), ctypes.byref(w))
    return m.value, e.value, w.value


# ---- MLP reference ---------------------------------------------------------

def mlp_param_count(nin: int, h1: int, h2: int) -> int:
    return lib().mlp_ref_param_count(nin, h1, h2)


def mlp_forward(p, nin, h1, h2, x) -> np.ndarray:
    p = np.ascontiguousarray(p, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    B = x.shape[0]
    y = np.empty(B)
    lib().mlp_ref_forward(_p(p), nin, h1, h2, B, _p(x), _p(y))
    return y


def mlp_loss(p, nin, h1, h2, x, t) -> float:
    p = np.ascontiguousarray(p, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    return lib().mlp_ref_loss(_p(p), nin, h1, h2, x.shape[0], _p(x), _p(t))


def mlp_grad(p, nin, h1, h2, x, t):
    p = np.ascontiguousarray(p, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    g = np.empty_like(p)
    loss = lib().mlp_ref_grad(_p(p), nin, h1, h2, x.shape[0], _p(x), _p(t), _p(g))
    return loss, g


def mlp_adam(p, g, m, v, lr, b1, b2, eps, t):
    for a in (p, m, v):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    g = np.ascontiguousarray(g, dtype=np.float64)
    lib().mlp_ref_adam(_p(p), _p(g), _p(m), _p(v), p.size, lr, b1, b2, eps, t)
