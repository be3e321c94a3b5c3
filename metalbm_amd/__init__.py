"""metalbm_amd -- MI355X-native fused collide-and-stream-pull step of metaLBM.

Host-side mirror of the reference's operator vocabulary over the C ABI of
libmetalbm_amd.so (include/metalbm_amd.h):

* enums with the reference's names and integer values
  (include/metaLBM/Options.h:12-41): ``LatticeType``, ``CollisionType``,
  ``ForcingSchemeType``, ``ForceType``, ``EquilibriumType``;
* ``Config`` -- the run parameters the reference takes from its compile-time
  input file (src/Input_prod.in:9-79) plus storage/compute precision;
* ``Algorithm`` -- one periodic domain on the GPU; ``iterate(n)`` advances n
  fused steps (Algorithm.h:326-358, Routine.h:121-147), ``store_fields``
  returns rho, u_hydro, alpha, force as storeFields writes them
  (Algorithm.h:150-194).

Errors raise :class:`LBMError` (the reference prints and exits,
Commons.h:6-38).  All numerics run in the HIP library; this module only moves
numpy arrays across the ABI.
"""
from __future__ import annotations

import ctypes
import enum
import os
from dataclasses import dataclass

import numpy as np

from .capi import LBMError, LbmConfigC, LbmElbmStatsC, check, lib

__all__ = [
    "LatticeType", "CollisionType", "ForcingSchemeType", "ForceType",
    "EquilibriumType", "Precision", "AlphaMode", "KernelShape", "Config", "Algorithm",
    "LBMError", "lattice_q", "lattice_d", "kernel_shape",
]


class LatticeType(enum.IntEnum):
    D2Q9 = 3
    D3Q19 = 9
    D3Q27 = 10


class CollisionType(enum.IntEnum):
    BGK = 1
    ELBM = 2


class ForcingSchemeType(enum.IntEnum):
    None_ = 1
    Guo = 2
    ShanChen = 3
    ExactDifferenceMethod = 4


class ForceType(enum.IntEnum):
    None_ = 0
    Constant = 4
    Sinusoidal = 5
    Kolmogorov = 6
    ConstantShell = 7    # spectral force field, set once (Force.h:296-395)
    EnergyRemoval = 8    # -A (rho u) on a shell, recomputed each step (Force.h:398-541)
    Turbulent2D = 9      # ConstantShell + EnergyRemoval (Force.h:544-600)


class EquilibriumType(enum.IntEnum):
    Exact = 1
    TruncationMa3 = 3


class Precision(enum.IntEnum):
    F64 = 0
    F32 = 1
    F16 = 2


class KernelShape(enum.IntEnum):
    AUTO = 0
    CELL = 1
    VECTOR = 2


class AlphaMode(enum.IntEnum):
    NEWTON = 0          # reference Newton (Collision.h:328-375)
    LEARNED = 1         # MLP start + one certified Newton correction, else NEWTON
    WARM_RESIDUAL = 2   # the same from the stored alpha (no network)
    NEWTON_RESIDUAL = 3  # plain Newton from the stored alpha until |G| <= residual_tol


_QD = {LatticeType.D2Q9: (9, 2), LatticeType.D3Q19: (19, 3), LatticeType.D3Q27: (27, 3)}


def kernel_shape(config: "Config"):
    """(KernelShape, V) that lbm_create would choose for ``config``; host-only."""
    shape, vec = ctypes.c_int(), ctypes.c_int()
    c = config.to_c()
    check(lib().lbm_kernel_shape(ctypes.byref(c), ctypes.byref(shape), ctypes.byref(vec)), None)
    return KernelShape(shape.value), vec.value


def lattice_q(lattice) -> int:
    return _QD[LatticeType(lattice)][0]


def lattice_d(lattice) -> int:
    return _QD[LatticeType(lattice)][1]


@dataclass
class Config:
    lattice: int = LatticeType.D2Q9
    collision: int = CollisionType.BGK
    forcing_scheme: int = ForcingSchemeType.Guo
    force: int = ForceType.None_
    equilibrium: int = EquilibriumType.TruncationMa3
    nx: int = 16
    ny: int = 16
    nz: int = 1
    tau: float = 0.6
    amplitude: tuple = (0.0, 0.0, 0.0)
    wavelength: tuple = (1.0, 1.0, 1.0)
    storage: int = Precision.F64
    shifted: bool = False
    compute: int = Precision.F64
    alpha_mode: int = AlphaMode.NEWTON
    residual_tol: float = 1e-10
    kernel: int = KernelShape.AUTO
    # spectral forces (Input_prod_2d.in:49-57 names)
    force_kmin: int = 1
    force_kmax: int = 2
    removal_amplitude: tuple = (1e-6, 1e-6, 1e-6)
    removal_kmin: int = 1
    removal_kmax: int = 2

    def to_c(self) -> LbmConfigC:
        c = LbmConfigC()
        for name in ("lattice", "collision", "forcing_scheme", "force",
                     "equilibrium", "nx", "ny", "nz", "storage", "compute",
                     "alpha_mode", "kernel", "force_kmin", "force_kmax",
                     "removal_kmin", "removal_kmax"):
            setattr(c, name, int(getattr(self, name)))
        for d in range(3):
            c.removal_amplitude[d] = float(self.removal_amplitude[d])
        c.shifted = 1 if self.shifted else 0
        c.tau = float(self.tau)
        c.residual_tol = float(self.residual_tol)
        for d in range(3):
            c.amplitude[d] = float(self.amplitude[d])
            c.wavelength[d] = float(self.wavelength[d])
        return c

    @property
    def q(self) -> int:
        return lattice_q(self.lattice)

    @property
    def d(self) -> int:
        return lattice_d(self.lattice)

    @property
    def cells(self) -> int:
        return self.nx * self.ny * self.nz


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f64(a, shape) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.size != int(np.prod(shape)):
        raise LBMError(-1, f"array of size {a.size} where {shape} was expected")
    return a.reshape(shape)


class Algorithm:
    """One periodic LBM domain resident on the current GPU."""

    def __init__(self, config: Config):
        self.config = config
        self._h = ctypes.c_void_p()
        c = config.to_c()
        check(lib().lbm_create(ctypes.byref(c), ctypes.byref(self._h)), None)
        self.N, self.Q, self.D = config.cells, config.q, config.d

    # -- lifetime -------------------------------------------------------------
    def close(self) -> None:
        if self._h:
            lib().lbm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        check(rc, self._h)

    # -- initial state ---------------------------------------------------------
    def init_equilibrium(self, rho, u) -> None:
        """f = feq(rho, u) (Initialize.h:106-117); rho[N], u[D, N]."""
        r = _f64(rho, (self.N,))
        v = _f64(u, (self.D, self.N))
        self._check(lib().lbm_init_equilibrium(self._h, _dp(r), _dp(v)))

    def init_noise(self, rho0: float, amplitude: float, seed: int) -> None:
        self._check(lib().lbm_init_noise(self._h, rho0, amplitude, seed))

    def init_taylor_green(self, rho0: float, u0: float) -> None:
        self._check(lib().lbm_init_taylor_green(self._h, rho0, u0))

    def init_spectral(self, rho0: float, urms: float, kmax: int = 16, kpeak: float = 4.0,
                      seed: int = 42) -> None:
        """f = feq(rho0, u0), u0 solenoidal with random phases and shell
        spectrum s^4 exp(-2 (s/kpeak)^2), s = 1..kmax, per-component rms urms."""
        self._check(lib().lbm_init_spectral(self._h, rho0, urms, kmax, kpeak, seed))

    # -- populations / fields --------------------------------------------------
    def set_populations(self, f) -> None:
        a = _f64(f, (self.Q, self.N))
        self._check(lib().lbm_set_populations(self._h, _dp(a)))

    def populations(self) -> np.ndarray:
        f = np.empty((self.Q, self.N))
        self._check(lib().lbm_get_populations(self._h, _dp(f)))
        return f

    def set_alpha(self, alpha) -> None:
        a = _f64(alpha, (self.N,))
        self._check(lib().lbm_set_alpha(self._h, _dp(a)))

    def alpha(self) -> np.ndarray:
        a = np.empty(self.N)
        self._check(lib().lbm_get_alpha(self._h, _dp(a)))
        return a

    def iterate(self, nsteps: int = 1, store: bool = False) -> None:
        """Advance nsteps fused steps; store=True writes the output fields on
        the last one (asynchronous, like a kernel launch)."""
        self._check(lib().lbm_step(self._h, int(nsteps), 1 if store else 0))

    def synchronize(self) -> None:
        self._check(lib().lbm_synchronize(self._h))

    def store_fields(self):
        """(rho[N], u_hydro[D, N], alpha[N], force[D, N]) of the last storing step."""
        rho = np.empty(self.N)
        u = np.empty((self.D, self.N))
        al = np.empty(self.N)
        F = np.empty((self.D, self.N))
        self._check(lib().lbm_get_fields(self._h, _dp(rho), _dp(u), _dp(al), _dp(F)))
        return rho, u, al, F

    def reduce(self):
        """(mass, energy, enstrophy) in FP64 from the stored fields."""
        m, e, w = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(lib().lbm_reduce(self._h, ctypes.byref(m), ctypes.byref(e),
                                     ctypes.byref(w)))
        return m.value, e.value, w.value

    def bandwidth_probe(self, kind: int, nsteps: int = 1) -> None:
        """nsteps launches of bandwidth probe `kind` (LBM_PROBE_*: 0 pull copy,
        1 read, 2 write, 3 contiguous copy); overwrites the populations."""
        self._check(lib().lbm_bandwidth_probe(self._h, int(kind), int(nsteps)))

    def reduce_noalloc(self):
        """reduce() through lbm_reduce_noalloc: no device allocation (static
        scratch of the library; grids up to 2^29 cells)."""
        m, e, w = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(lib().lbm_reduce_noalloc(self._h, ctypes.byref(m), ctypes.byref(e),
                                             ctypes.byref(w)))
        return m.value, e.value, w.value

    def elbm_stats(self) -> dict:
        s = LbmElbmStatsC()
        self._check(lib().lbm_get_elbm_stats(self._h, ctypes.byref(s)))
        return {k: int(getattr(s, k)) for k, _ in LbmElbmStatsC._fields_}

    def reset_elbm_stats(self) -> None:
        self._check(lib().lbm_reset_elbm_stats(self._h))

    def set_stream(self, stream_ptr: int | None) -> None:
        self._check(lib().lbm_set_stream(self._h, ctypes.c_void_p(stream_ptr or 0)))

    def device_populations(self):
        p = ctypes.c_void_p()
        e = ctypes.c_int()
        self._check(lib().lbm_device_populations(self._h, ctypes.byref(p), ctypes.byref(e)))
        return p.value, e.value

    @property
    def steps_done(self) -> int:
        return int(lib().lbm_steps_done(self._h))

    # -- learned alpha ---------------------------------------------------------
    def mlp_shape(self):
        """(nin, h1, h2, nparams) of the alpha predictor for this lattice."""
        v = [ctypes.c_int() for _ in range(4)]
        self._check(lib().lbm_mlp_shape(self._h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def mlp_init(self, seed: int) -> None:
        self._check(lib().lbm_mlp_init(self._h, seed))

    def mlp_params(self) -> np.ndarray:
        p = np.empty(self.mlp_shape()[3], dtype=np.float32)
        self._check(lib().lbm_mlp_get_params(self._h, _fp(p)))
        return p

    def set_mlp_params(self, p) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != self.mlp_shape()[3]:
            raise LBMError(-1, "wrong parameter count")
        self._check(lib().lbm_mlp_set_params(self._h, _fp(a)))

    def set_normalization(self, in_mean, in_scale, out_mean: float, out_scale: float) -> None:
        m = np.ascontiguousarray(in_mean, dtype=np.float32)
        s = np.ascontiguousarray(in_scale, dtype=np.float32)
        self._check(lib().lbm_mlp_set_normalization(self._h, _fp(m), _fp(s), out_mean, out_scale))

    def normalization(self):
        nin = self.mlp_shape()[0]
        m = np.empty(nin, dtype=np.float32)
        s = np.empty(nin, dtype=np.float32)
        om, os_ = ctypes.c_double(), ctypes.c_double()
        self._check(lib().lbm_mlp_get_normalization(self._h, _fp(m), _fp(s), ctypes.byref(om),
                                                    ctypes.byref(os_)))
        return m, s, om.value, os_.value

    def mlp_gradient(self, x, alpha):
        """(grad[nparams], loss) over samples x[n, nin] with targets alpha[n]."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        a = np.ascontiguousarray(alpha, dtype=np.float64)
        g = np.empty(self.mlp_shape()[3], dtype=np.float32)
        loss = ctypes.c_double()
        self._check(lib().lbm_mlp_gradient(self._h, _fp(x), _dp(a), int(a.size), _fp(g),
                                           ctypes.byref(loss)))
        return g, loss.value

    def mlp_predict(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        n = x.shape[0]
        out = np.empty(n)
        self._check(lib().lbm_mlp_predict(self._h, _fp(x), n, _dp(out)))
        return out

    def sample_log_enable(self, capacity: int, fraction: float = 1.0, salt: int = 0) -> None:
        self._check(lib().lbm_sample_log_enable(self._h, capacity, fraction, salt))

    def sample_log_disable(self) -> None:
        self._check(lib().lbm_sample_log_disable(self._h))

    def sample_log(self):
        """(features[n, nin], alpha[n]) logged so far."""
        n = ctypes.c_int64()
        self._check(lib().lbm_sample_log_count(self._h, ctypes.byref(n)))
        nin = self.mlp_shape()[0]
        x = np.empty((n.value, nin), dtype=np.float32)
        a = np.empty(n.value)
        if n.value:
            self._check(lib().lbm_sample_log_copy(self._h, _fp(x), _dp(a), n.value))
        return x, a

    def mlp_train(self, iterations: int, batch: int, lr: float, seed: int = 0):
        """Adam on the logged samples; returns (first loss, last loss)."""
        l0, l1 = ctypes.c_double(), ctypes.c_double()
        self._check(lib().lbm_mlp_train(self._h, iterations, batch, lr, seed, ctypes.byref(l0),
                                        ctypes.byref(l1)))
        return l0.value, l1.value

    def set_alpha_mode(self, mode, residual_tol: float = 1e-10) -> None:
        self._check(lib().lbm_set_alpha_mode(self._h, int(mode), residual_tol))

    # -- HDF5 files in the reference layout -------------------------------------
    # File names follow Writer::getFileName (Writer.h:44-50): <folder>/field-<it>.h5
    # with its field-<it>.xmf sidecar, and <folder>/distribution-<it>.h5.
    def set_field_output(self, alpha: bool = True, force: bool = True,
                         vorticity: bool = True) -> None:
        """writeAlpha / writeForce / writeVorticity of the field files
        (Input_prod.in:68-72; all on by default)."""
        self._check(lib().lbm_set_field_output(self._h, int(alpha), int(force), int(vorticity)))

    def write_fields_hdf5(self, path: str, xdmf_path: str | None = None) -> None:
        """Fields of the last storing step as FP64 datasets (density,
        velocityX.., alpha for ELBM, forceX.. when forced)."""
        self._check(lib().lbm_write_fields_hdf5(
            self._h, os.fsencode(path), os.fsencode(xdmf_path) if xdmf_path else None))

    def write_distribution_hdf5(self, path: str) -> None:
        self._check(lib().lbm_write_distribution_hdf5(self._h, os.fsencode(path)))

    def read_distribution_hdf5(self, path: str) -> None:
        self._check(lib().lbm_read_distribution_hdf5(self._h, os.fsencode(path)))

    def write_fields(self, folder: str, iteration: int) -> str:
        """Routine::writeFields (Routine.h:202-210) into `folder`; returns the .h5 path."""
        os.makedirs(folder, exist_ok=True)
        path = os.path.join(folder, f"field-{int(iteration)}.h5")
        self.write_fields_hdf5(path, os.path.join(folder, f"field-{int(iteration)}.xmf"))
        return path

    def write_distribution(self, folder: str, iteration: int) -> str:
        """The distribution backup of Routine::writeFields (Routine.h:212-216)."""
        os.makedirs(folder, exist_ok=True)
        path = os.path.join(folder, f"distribution-{int(iteration)}.h5")
        self.write_distribution_hdf5(path)
        return path

    def read_distribution(self, folder: str, iteration: int) -> None:
        """Restart from <folder>/distribution-<iteration>.h5 (Initialize.h:119-123)."""
        self.read_distribution_hdf5(os.path.join(folder, f"distribution-{int(iteration)}.h5"))

    # -- spectral diagnostics of the stored fields ------------------------------
    _SCALING = {"reference": 0, "physical": 1}

    def spectral_shells(self) -> int:
        n = ctypes.c_int()
        self._check(lib().lbm_spectral_shells(self._h, ctypes.byref(n)))
        return n.value

    def vorticity(self, scaling: str = "reference", field: bool = True):
        """(omega[2D-3, N], enstrophy) of the stored u_hydro, computed
        spectrally (Curl, Transformer.h:114-295; TotalEnstrophy,
        Analysis.h:68-98) in the "reference" or "physical" scaling;
        field=False returns (None, enstrophy) without copying omega."""
        C = 3 if self.D == 3 else 1
        w = np.empty((C, self.N)) if field else None
        e = ctypes.c_double()
        self._check(lib().lbm_vorticity(self._h, self._SCALING[scaling],
                                        _dp(w) if field else None, ctypes.byref(e)))
        return w, e.value

    def spectra(self, scaling: str = "reference"):
        """(energy[nshells], forcing[nshells]) shell spectra of u_hydro and of
        the force field (PowerSpectra, AnalysisList.h:98-202)."""
        ns = self.spectral_shells()
        e = np.empty(max(ns, 1))
        f = np.empty(max(ns, 1))
        self._check(lib().lbm_spectra(self._h, self._SCALING[scaling], _dp(e), _dp(f)))
        return e[:ns], f[:ns]

    def handle_config(self) -> LbmConfigC:
        """The handle's configuration as the C struct (alpha mode as last set)."""
        c = LbmConfigC()
        self._check(lib().lbm_get_config(self._h, ctypes.byref(c)))
        return c

