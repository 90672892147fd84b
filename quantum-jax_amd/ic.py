"""Seeded initial conditions of the benchmark workloads (host FP64).

psi: Re, Im ~ N(0,1) i.i.d. per cell (Re drawn first, numpy PCG64), multiplied
in k-space by exp(-|k|^2/(2 sigma_k^2)) with sigma_k = 16 pi, rescaled to
mean |psi|^2 = 1.  Stars: uniform FP64 positions in the box, N(0, 0.05^2)
velocities, equal masses totalling 0.1.  Gas: rho = 0.2 (1 + 0.1 g) with g
the real part of the filtered field scaled to zero mean and unit variance,
clipped below at 0.01; v = 0.

These seeded constructions define the workloads that the tests and the
benchmark share.  The reference's own ICs are below them: random phases
sorted by |k| (quantum-fluid.py:509-522) and the Gaussian columns with
uniform stars (quantum-stars.py:376-428).  The cosmological variant's HDF5
initial condition (quantum-cosmo.py:242-244) is ``hdf5_psi``, read by the
package's own HDF5 reader (``hdf5.py``).
"""

from __future__ import annotations

import numpy as np
import scipy.fft as sfft

from .hdf5 import load_psi as hdf5_psi  # noqa: F401


def kvec(n: int, L: float) -> np.ndarray:
    """FFT-order wavenumbers, Nyquist negative (quantum-stars.py:142-148)."""
    return np.fft.ifftshift(2.0 * np.pi / L * np.arange(-n / 2, n / 2))


def k_squared(shape, L=(1.0, 1.0, 1.0)):
    kx = kvec(shape[0], L[0])[:, None, None]
    ky = kvec(shape[1], L[1])[None, :, None]
    kz = kvec(shape[2], L[2])[None, None, :]
    return kx ** 2 + ky ** 2 + kz ** 2


def _workers():
    import os
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def filtered_gaussian_field(shape, seed: int, L=(1.0, 1.0, 1.0), sigma_k: float = 16.0 * np.pi):
    rng = np.random.default_rng(seed)
    re = rng.standard_normal(shape)
    im = rng.standard_normal(shape)
    f = sfft.fftn(re + 1j * im, workers=_workers())
    f *= np.exp(-k_squared(shape, L) / (2.0 * sigma_k ** 2))
    return sfft.ifftn(f, workers=_workers())


def filtered_gaussian_psi(shape, seed: int, L=(1.0, 1.0, 1.0), mean_rho: float = 1.0,
                          sigma_k: float = 16.0 * np.pi):
    f = filtered_gaussian_field(shape, seed, L, sigma_k)
    return f * np.sqrt(mean_rho / np.mean(np.abs(f) ** 2))


def star_ic(n_p: int, seed: int, L=(1.0, 1.0, 1.0), total_mass: float = 0.1,
            sigma_v: float = 0.05):
    rng = np.random.default_rng(seed)
    pos = rng.random((n_p, 3)) * np.asarray(L, dtype=np.float64)
    vel = rng.normal(0.0, sigma_v, (n_p, 3))
    return pos, vel, (total_mass / n_p if n_p else 0.0)


def gas_ic(shape, seed: int, L=(1.0, 1.0, 1.0), base: float = 0.2, amp: float = 0.1,
           floor: float = 0.01, sigma_k: float = 16.0 * np.pi):
    f = np.real(filtered_gaussian_field(shape, seed, L, sigma_k))
    f = (f - f.mean()) / f.std()
    rho = np.maximum(base * (1.0 + amp * f), floor)
    z = np.zeros(shape)
    return rho, z.copy(), z.copy(), z.copy()


# ---------------------------------------------------------------------------
# The reference's own initial conditions
# ---------------------------------------------------------------------------
def k_squared_f32(shape, L=(1.0, 1.0, 1.0)):
    """k^2 as the reference computes it at its default precision (x64 off,
    quantum-fluid.py:50-51,128-134): k = float32(2 pi / L) * float32(i - n/2),
    ifftshifted, and k^2 = (kx*kx + ky*ky) + kz*kz, every operation rounded
    to float32."""
    ks = []
    for n, Li in zip(shape, L):
        c = np.float32(2.0 * np.pi / Li)
        ks.append(np.fft.ifftshift(c * np.arange(-n / 2, n / 2, dtype=np.float32)))
    kx, ky, kz = np.meshgrid(*ks, indexing="ij")
    return (kx * kx + ky * ky) + kz * kz


def random_phase_psi(shape, L, sigma: float, m_per_hbar: float, mean_rho: float,
                     seed: int = 17, k2_precision: str = "float32"):
    """quantum-fluid.py:509-522: unit-modulus random phases assigned in
    increasing |k| order (stable argsort of k^2, so the lowest modes take the
    first numbers of NumPy's legacy seeded generator), amplitude
    sqrt(exp(-k^2 / (2 sigma^2 m_per_hbar^2))), inverse FFT, and mean
    |psi|^2 = mean_rho (the reference's frac_dm * rho_bar).

    k2_precision sets the sort key: "float32" is k^2 rounded as the
    reference computes it by default (x64 off, k_squared_f32), "float64" as
    it computes it with jax_enable_x64.  Modes of one |k| shell tie in exact
    arithmetic, and the two roundings break those ties differently (at 32^3
    the orders part from the 98th mode on), so each mode gets a different
    random phase.  The amplitudes and the normalisation are FP64 either way."""
    if k2_precision not in ("float32", "float64"):
        raise ValueError("k2_precision must be 'float32' or 'float64'")
    k2 = k_squared(shape, L)
    key = k_squared_f32(shape, L) if k2_precision == "float32" else k2
    rs = np.random.RandomState(seed)  # np.random.seed(seed); np.random.rand
    order = np.argsort(key.ravel(), kind="stable")
    psi = np.zeros(k2.size, dtype=complex)
    psi[order] = np.exp(1.0j * 2.0 * np.pi * rs.rand(k2.size))
    psi = psi.reshape(k2.shape) * np.sqrt(np.exp(-k2 / (2.0 * sigma ** 2 * m_per_hbar ** 2)))
    psi = np.fft.ifftn(psi)
    return psi * np.sqrt(mean_rho / np.mean(np.abs(psi) ** 2))


# (weight, x / Lx, y / Ly) of the Gaussian columns of quantum-stars.py:379-420
STAR_BLOBS = ((2.0, 0.5, 0.4), (1.5, 0.6, 0.5), (1.0, 0.4, 0.6), (1.0, 0.6, 0.4),
              (1.0, 0.6, 0.6), (1.0, 0.6, 0.4), (1.0, 0.5, 0.4), (1.0, 0.5, 0.4))


def blob_psi(shape, L, mean_rho: float, amp: float = 100.0, sigma: float = 4.0,
             floor: float = 10.0, blobs=STAR_BLOBS):
    """quantum-stars.py:376-423: rho = floor + sum of Gaussian columns (in x
    and y, uniform in z) of weight w amp / (sigma^3 2 pi), rescaled to mean
    mean_rho (frac_dm * rho_bar); psi = sqrt(rho)."""
    axes = [np.linspace(0.5 * Li / n, Li - 0.5 * Li / n, n) for n, Li in zip(shape, L)]
    X, Y, _ = np.meshgrid(*axes, indexing="ij")
    rho = np.full(shape, float(floor))
    for w, cx, cy in blobs:
        rho = rho + (w * amp * np.exp(-((X - cx * L[0]) ** 2 + (Y - cy * L[1]) ** 2) / 2.0
                                      / sigma ** 2) / (sigma ** 3 * np.sqrt(2.0 * np.pi) ** 2))
    rho *= mean_rho / np.mean(rho)
    return np.sqrt(rho) + 0.0j


def uniform_star_ic(n_s: int, L, seed: int = 17):
    """quantum-stars.py:425-428: positions uniform in the box and velocities
    uniform in [-1, 1] from NumPy's legacy generator seeded with seed."""
    rs = np.random.RandomState(seed)
    pos = rs.uniform(0.0, 1.0, (n_s, 3)) * np.asarray(L, dtype=np.float64)
    vel = rs.uniform(-1.0, 1.0, (n_s, 3))
    return pos, vel
