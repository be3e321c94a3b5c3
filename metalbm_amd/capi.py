"""ctypes binding of libmetalbm_amd.so (include/metalbm_amd.h).

Loads the in-tree library built by ``python -m metalbm_amd.build`` and fails
loudly if it is missing: there is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmetalbm_amd.so")

LBM_OK, LBM_E_INVALID, LBM_E_HIP, LBM_E_NOMEM, LBM_E_STATE = 0, -1, -2, -3, -4
LBM_SPECTRAL_REFERENCE, LBM_SPECTRAL_PHYSICAL = 0, 1

# Every entry point declared in include/metalbm_amd.h.
EXPORTS = (
    "lbm_abi_version", "lbm_config_default", "lbm_create", "lbm_destroy",
    "lbm_last_error", "lbm_sizes", "lbm_init_equilibrium", "lbm_init_noise",
    "lbm_init_taylor_green", "lbm_init_spectral", "lbm_set_populations", "lbm_get_populations",
    "lbm_set_alpha", "lbm_get_alpha", "lbm_step", "lbm_get_fields", "lbm_reduce",
    "lbm_get_elbm_stats", "lbm_reset_elbm_stats", "lbm_set_stream",
    "lbm_synchronize", "lbm_device_populations", "lbm_steps_done",
    "lbm_mlp_shape", "lbm_mlp_init", "lbm_mlp_set_params", "lbm_mlp_get_params",
    "lbm_mlp_set_normalization", "lbm_mlp_get_normalization", "lbm_mlp_gradient",
    "lbm_mlp_predict", "lbm_sample_log_enable", "lbm_sample_log_disable",
    "lbm_sample_log_count", "lbm_sample_log_copy", "lbm_mlp_train", "lbm_set_alpha_mode",
    "lbm_get_config", "lbm_write_fields_hdf5", "lbm_write_distribution_hdf5",
    "lbm_read_distribution_hdf5", "lbm_spectral_shells", "lbm_vorticity", "lbm_spectra",
    "lbm_kernel_shape", "lbm_reduce_noalloc", "lbm_bandwidth_probe", "lbm_set_field_output",
)


class LbmConfigC(ctypes.Structure):
    _fields_ = [
        ("lattice", ctypes.c_int), ("collision", ctypes.c_int),
        ("forcing_scheme", ctypes.c_int), ("force", ctypes.c_int),
        ("equilibrium", ctypes.c_int),
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("tau", ctypes.c_double),
        ("amplitude", ctypes.c_double * 3), ("wavelength", ctypes.c_double * 3),
        ("storage", ctypes.c_int), ("shifted", ctypes.c_int),
        ("compute", ctypes.c_int), ("alpha_mode", ctypes.c_int),
        ("residual_tol", ctypes.c_double), ("kernel", ctypes.c_int),
        ("force_kmin", ctypes.c_int), ("force_kmax", ctypes.c_int),
        ("removal_amplitude", ctypes.c_double * 3),
        ("removal_kmin", ctypes.c_int), ("removal_kmax", ctypes.c_int),
    ]


class LbmElbmStatsC(ctypes.Structure):
    _fields_ = [("small", ctypes.c_int64), ("capped", ctypes.c_int64),
                ("newton", ctypes.c_int64), ("fallback", ctypes.c_int64),
                ("newton_iterations", ctypes.c_int64),
                ("learned_accepted", ctypes.c_int64)]


class LBMError(RuntimeError):
    """A failed C-ABI call: ``code`` is the LBM_E_* value."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m metalbm_amd.build` "
            "(the HIP library is required; there is no fallback)")
    L = ctypes.CDLL(LIB_PATH)
    dp = ctypes.POINTER(ctypes.c_double)
    H = ctypes.c_void_p
    L.lbm_abi_version.restype = ctypes.c_int
    L.lbm_config_default.argtypes = [ctypes.POINTER(LbmConfigC)]
    L.lbm_create.argtypes = [ctypes.POINTER(LbmConfigC), ctypes.POINTER(H)]
    L.lbm_destroy.argtypes = [H]
    L.lbm_last_error.argtypes = [H]
    L.lbm_last_error.restype = ctypes.c_char_p
    L.lbm_sizes.argtypes = [H, ctypes.POINTER(ctypes.c_int64),
                            ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.lbm_init_equilibrium.argtypes = [H, dp, dp]
    L.lbm_init_noise.argtypes = [H, ctypes.c_double, ctypes.c_double, ctypes.c_uint64]
    L.lbm_init_taylor_green.argtypes = [H, ctypes.c_double, ctypes.c_double]
    L.lbm_init_spectral.argtypes = [H, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_uint64]
    L.lbm_set_populations.argtypes = [H, dp]
    L.lbm_get_populations.argtypes = [H, dp]
    L.lbm_set_alpha.argtypes = [H, dp]
    L.lbm_get_alpha.argtypes = [H, dp]
    L.lbm_step.argtypes = [H, ctypes.c_int, ctypes.c_int]
    L.lbm_get_fields.argtypes = [H, dp, dp, dp, dp]
    L.lbm_reduce.argtypes = [H, dp, dp, dp]
    L.lbm_reduce_noalloc.argtypes = [H, dp, dp, dp]
    L.lbm_bandwidth_probe.argtypes = [H, ctypes.c_int, ctypes.c_int]
    L.lbm_get_elbm_stats.argtypes = [H, ctypes.POINTER(LbmElbmStatsC)]
    L.lbm_reset_elbm_stats.argtypes = [H]
    L.lbm_set_stream.argtypes = [H, ctypes.c_void_p]
    L.lbm_synchronize.argtypes = [H]
    L.lbm_device_populations.argtypes = [H, ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.POINTER(ctypes.c_int)]
    L.lbm_steps_done.argtypes = [H]
    L.lbm_steps_done.restype = ctypes.c_int64
    ip = ctypes.POINTER(ctypes.c_int)
    fp = ctypes.POINTER(ctypes.c_float)
    L.lbm_mlp_shape.argtypes = [H, ip, ip, ip, ip]
    L.lbm_mlp_init.argtypes = [H, ctypes.c_uint64]
    L.lbm_mlp_set_params.argtypes = [H, fp]
    L.lbm_mlp_get_params.argtypes = [H, fp]
    L.lbm_mlp_set_normalization.argtypes = [H, fp, fp, ctypes.c_double, ctypes.c_double]
    L.lbm_mlp_get_normalization.argtypes = [H, fp, fp, dp, dp]
    L.lbm_mlp_gradient.argtypes = [H, fp, dp, ctypes.c_int, fp, dp]
    L.lbm_mlp_predict.argtypes = [H, fp, ctypes.c_int, dp]
    L.lbm_sample_log_enable.argtypes = [H, ctypes.c_int64, ctypes.c_double, ctypes.c_uint32]
    L.lbm_sample_log_disable.argtypes = [H]
    L.lbm_sample_log_count.argtypes = [H, ctypes.POINTER(ctypes.c_int64)]
    L.lbm_sample_log_copy.argtypes = [H, fp, dp, ctypes.c_int64]
    L.lbm_mlp_train.argtypes = [H, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64,
                                dp, dp]
    L.lbm_set_alpha_mode.argtypes = [H, ctypes.c_int, ctypes.c_double]
    L.lbm_get_config.argtypes = [H, ctypes.POINTER(LbmConfigC)]
    L.lbm_kernel_shape.argtypes = [ctypes.POINTER(LbmConfigC), ip, ip]
    cp = ctypes.c_char_p
    L.lbm_write_fields_hdf5.argtypes = [H, cp, cp]
    L.lbm_set_field_output.argtypes = [H, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.lbm_write_distribution_hdf5.argtypes = [H, cp]
    L.lbm_read_distribution_hdf5.argtypes = [H, cp]
    L.lbm_spectral_shells.argtypes = [H, ip]
    L.lbm_vorticity.argtypes = [H, ctypes.c_int, dp, dp]
    L.lbm_spectra.argtypes = [H, ctypes.c_int, dp, dp]
    _lib = L
    return L


def check(rc: int, handle=None) -> None:
    if rc != LBM_OK:
        msg = lib().lbm_last_error(handle)
        raise LBMError(rc, msg.decode() if msg else "unknown error")
