"""numpy restatement of the reference's spectral diagnostics.

TEST INFRASTRUCTURE ONLY.  Only tests/ may import this module, and only as
the checker of the library's ``lbm_vorticity`` / ``lbm_spectra``.

Follows, on the FFTW r2c half spectrum (numpy's ``rfftn`` has the same
layout: last axis 0..n/2):

* ``Curl<double, CPU, OneD, 2, 2>::executeFourier`` (Transformer.h:146-172):
  omega_hat = i (kx uy_hat - ky ux_hat) with integer wavenumbers
  (i <= n/2 -> i, else i - n), imaginary part set to zero;
* ``Curl<..., 3, 3>::executeFourier`` (Transformer.h:216-275): the three
  components of i k x u_hat;
* the backward transform divides by N (Transformer.h:100-111) and
  ``Curl::normalize`` by N again (Transformer.h:179-185, 284-294);
* ``TotalEnstrophy`` (Analysis.h:68-98, 27): (1/N) sum 1/2 |omega|^2;
* ``PowerSpectra::operator()`` (Analysis.h:146-165): per mode
  c * sum_d |u_hat_d|^2 added to shell ``int(sqrt(|k|^2))`` when below
  max(n)/2 (AnalysisList.h:143-161, FourierDomain.h:76-78), c = 1/2 on the
  k_last = 0 plane; ``normalize`` divides the energy spectrum by N, the
  forcing spectrum is left unnormalised (AnalysisList.h:195).

``scaling="physical"`` is this library's physical variant (no reference
counterpart): wavenumbers 2 pi k / n, one 1/N, the full complex 2D curl,
nearest-integer shells, c = 1/2 also on an even last axis' k = n/2 plane,
both spectra divided by N^2, the derivative zero at every Nyquist index.
The reference scaling keeps k = +n/2 at Nyquist indices except on the
self-paired planes k_last = 0 and k_last = n_last/2, where the inverse
transform returns the Hermitian part of the spectrum, in which that
derivative vanishes (zero there).

Fields are ``u[D, N]`` with cells ``(x * ny + y) * nz + z`` (2D: x * ny + y).
"""
from __future__ import annotations

import numpy as np

REFERENCE, PHYSICAL = "reference", "physical"


def _axes_k(shape):
    """Integer wavenumber grids over the half spectrum, and per axis the mask
    of its Nyquist index."""
    D = len(shape)
    ks, nyq = [], []
    for d, n in enumerate(shape):
        m = n // 2 + 1 if d == D - 1 else n
        i = np.arange(m)
        ks.append(np.where(i <= n // 2, i, i - n).astype(np.float64))
        nyq.append((i == n // 2) & (n % 2 == 0))
    return np.meshgrid(*ks, indexing="ij"), np.meshgrid(*nyq, indexing="ij")


def _derivative_k(shape, physical: bool):
    k, nyq = _axes_k(shape)
    n_last = shape[-1]
    l = np.arange(n_last // 2 + 1)
    self_paired = (l == 0) | ((l == n_last // 2) & (n_last % 2 == 0))
    out = []
    for d, n in enumerate(shape):
        zero = nyq[d] if physical else (nyq[d] & self_paired)
        kd = np.where(zero, 0.0, k[d])
        out.append(kd * (2.0 * np.pi / n) if physical else kd)
    return out


def _hat(u, shape):
    D = len(shape)
    return [np.fft.rfftn(np.asarray(u[d], dtype=np.float64).reshape(shape)) for d in range(D)]


def vorticity(u, shape, scaling=REFERENCE):
    """(omega[2D-3, N], enstrophy) of the velocity u[D, N] on a periodic grid."""
    shape = tuple(int(s) for s in shape)
    D = len(shape)
    N = int(np.prod(shape))
    phys = scaling == PHYSICAL
    k = _derivative_k(shape, phys)
    uh = _hat(u, shape)
    if D == 2:
        w = 1j * (k[0] * uh[1] - k[1] * uh[0])
        if not phys:
            w = w.real + 0j  # Transformer.h:161-165
        comps = [w]
    else:
        comps = [1j * (k[1] * uh[2] - k[2] * uh[1]),
                 1j * (k[2] * uh[0] - k[0] * uh[2]),
                 1j * (k[0] * uh[1] - k[1] * uh[0])]
    # np.fft.irfftn includes the backward transform's 1/N
    out = np.stack([np.fft.irfftn(c, s=shape, axes=tuple(range(D))).reshape(-1) for c in comps])
    if not phys:
        out = out / N  # Curl::normalize
    enstrophy = float(np.sum(0.5 * out * out) / N)
    return out, enstrophy


def shells(shape):
    return max(int(s) for s in shape) // 2


def spectrum(v, shape, scaling=REFERENCE, forcing=False):
    """Shell spectrum of the D-component field v[D, N]: the energy spectrum,
    or with ``forcing`` the reference's unnormalised forcing spectrum."""
    shape = tuple(int(s) for s in shape)
    D = len(shape)
    N = int(np.prod(shape))
    phys = scaling == PHYSICAL
    k, _ = _axes_k(shape)
    k2 = sum(kk * kk for kk in k)
    r = np.sqrt(k2)
    b = np.floor(r + 0.5).astype(np.int64) if phys else r.astype(np.int64)
    vh = _hat(v, shape)
    e = sum(h.real * h.real + h.imag * h.imag for h in vh)
    last = np.arange(shape[-1] // 2 + 1)
    c_last = np.where(last == 0, 0.5, 1.0)
    if phys and shape[-1] % 2 == 0:
        c_last = np.where(last == shape[-1] // 2, 0.5, c_last)
    e = e * c_last  # broadcasts over the last axis
    ns = shells(shape)
    keep = b < ns
    out = np.bincount(b[keep].reshape(-1), weights=e[keep].reshape(-1), minlength=ns)[:ns]
    if phys:
        return out / (N * float(N))
    return out if forcing else out / N


# -- spectral forces (Force.h:294-623) ----------------------------------------
# Shells are integer |k|^2 in [kmin^2, kmax^2] (Force.h:372, 463 kNormSquared).
# The reference's 3D mirror writes on the k_z = 0 plane use the y index
# n_y - 0 = n_y for k_y = 0 (Force.h:339-352), one x row past the intended
# mirror (and past the array for k_x = 1); restated here is the intended
# shell -- every half-spectrum mode of the shell -- which is what the 2D
# code (where that index is overwritten with 0) computes.

def _shell(shape, kmin, kmax):
    k, _ = _axes_k(shape)
    k2 = sum(kk * kk for kk in k)
    return (k2 >= kmin * kmin) & (k2 <= kmax * kmax)


def constant_shell_force(shape, amplitude, kmin, kmax):
    """ConstantShell::initTempArray + setForceArray (Force.h:327-393, 318-325):
    psi_hat_c = amplitude[c] on the shell, F = MakeIncompressible(psi)
    (Transformer.h:320-334 in 2D: F_hat = (i ky psi_hat, -i kx psi_hat); 3D:
    the Curl, Transformer.h:216-275) with integer wavenumbers, and the
    backward transform's 1/N.  Returns F[D, N]."""
    shape = tuple(int(s) for s in shape)
    D = len(shape)
    k = _derivative_k(shape, physical=False)
    mask = _shell(shape, kmin, kmax).astype(np.float64)
    if D == 2:
        psi = amplitude[0] * mask
        Fh = [1j * k[1] * psi, -1j * k[0] * psi]
    else:
        psi = [amplitude[c] * mask for c in range(3)]
        Fh = [1j * (k[1] * psi[2] - k[2] * psi[1]),
              1j * (k[2] * psi[0] - k[0] * psi[2]),
              1j * (k[0] * psi[1] - k[1] * psi[0])]
    ax = tuple(range(D))
    return np.stack([np.fft.irfftn(h, s=shape, axes=ax).reshape(-1) for h in Fh])


def energy_removal_force(rho, u, shape, amplitude, kmin, kmax):
    """EnergyRemoval::setForceArray (Force.h:416-497): F_d = -amplitude_d
    (rho u_d) restricted to the shell, through the forward transform, the
    shell product and the backward transform with its 1/N.  rho[N], u[D, N]
    are the stored fields.  Returns F[D, N]."""
    shape = tuple(int(s) for s in shape)
    D = len(shape)
    mask = _shell(shape, kmin, kmax)
    rho = np.asarray(rho, dtype=np.float64).reshape(-1)
    ax = tuple(range(D))
    out = []
    for d in range(D):
        ph = np.fft.rfftn((rho * np.asarray(u[d], dtype=np.float64).reshape(-1)).reshape(shape),
                          axes=ax)
        out.append(np.fft.irfftn(np.where(mask, -amplitude[d] * ph, 0.0), s=shape,
                                 axes=ax).reshape(-1))
    return np.stack(out)


def turbulent_force(rho, u, shape, amplitude, kmin, kmax, removal_amplitude, removal_kmin,
                    removal_kmax):
    """Turbulent2D::setForceArray (Force.h:565-588): injection (ConstantShell)
    plus removal (EnergyRemoval), summed in real space."""
    return (constant_shell_force(shape, amplitude, kmin, kmax)
            + energy_removal_force(rho, u, shape, removal_amplitude, removal_kmin, removal_kmax))
