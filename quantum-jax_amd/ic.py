"""Seeded initial conditions of the benchmark workloads (host FP64).

psi: Re, Im ~ N(0,1) i.i.d. per cell (Re drawn first, numpy PCG64), multiplied
in k-space by exp(-|k|^2/(2 sigma_k^2)) with sigma_k = 16 pi, rescaled to
mean |psi|^2 = 1.  Stars: uniform FP64 positions in the box, N(0, 0.05^2)
velocities, equal masses totalling 0.1.  Gas: rho = 0.2 (1 + 0.1 g) with g
the real part of the filtered field scaled to zero mean and unit variance,
clipped below at 0.01; v = 0.

The reference builds its own ICs differently (random phases sorted by |k|,
quantum-fluid.py:509-522); these seeded constructions define the workloads
that the tests and the benchmark share.
"""

from __future__ import annotations

import numpy as np
import scipy.fft as sfft


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
