"""The reference's initial-condition generators (host, CPU)."""

import importlib

import numpy as np
import pytest

ic = importlib.import_module("quantum-jax_amd.ic")


def test_random_phase_psi_amplitude_phases_and_order():
    shape, L, sigma, mph, mean_rho = (16, 16, 16), (1.0, 1.0, 1.0), 40.0, 0.5, 3.0
    psi = ic.random_phase_psi(shape, L, sigma, mph, mean_rho, seed=17)
    assert np.mean(np.abs(psi) ** 2) == pytest.approx(mean_rho, rel=1e-12)
    F = np.fft.fftn(psi)
    k2 = ic.k_squared(shape, L)
    amp = np.sqrt(np.exp(-k2 / (2 * sigma ** 2 * mph ** 2)))
    ratio = np.abs(F) / amp
    assert np.allclose(ratio, ratio.flat[0], rtol=1e-9)  # unit-modulus phases
    # the stream of the seeded legacy generator, in increasing-|k| order:
    # k = 0 first, then the six |k| = 2 pi modes in flat-index order
    u = np.random.RandomState(17).rand(7)
    assert np.angle(F.flat[0]) % (2 * np.pi) == pytest.approx(2 * np.pi * u[0], abs=1e-9)
    first_shell = np.flatnonzero(np.isclose(k2.ravel(), (2 * np.pi) ** 2))
    assert first_shell.size == 6
    ph = np.angle(F.ravel()[first_shell]) % (2 * np.pi)
    assert np.allclose(ph, 2 * np.pi * u[1:7], atol=1e-9)


def test_blob_psi_density_and_peak():
    shape, L = (32, 32, 8), (128.0, 128.0, 32.0)
    psi = ic.blob_psi(shape, L, mean_rho=5.0)
    rho = np.abs(psi) ** 2
    assert np.all(psi.imag == 0) and np.all(psi.real > 0)
    assert rho.mean() == pytest.approx(5.0, rel=1e-12)
    # uniform along z; the peak sits on the heaviest column, (0.5, 0.4)
    assert np.allclose(rho, rho[:, :, :1])
    i, j = np.unravel_index(np.argmax(rho[:, :, 0]), shape[:2])
    assert abs((i + 0.5) * 4.0 - 64.0) <= 4.0 and abs((j + 0.5) * 4.0 - 51.2) <= 4.0


def test_uniform_star_ic_stream():
    pos, vel = ic.uniform_star_ic(100, (2.0, 1.0, 3.0), seed=17)
    rs = np.random.RandomState(17)
    assert np.array_equal(pos, rs.uniform(0, 1, (100, 3)) * np.array([2.0, 1.0, 3.0]))
    assert np.array_equal(vel, rs.uniform(-1, 1, (100, 3)))
    assert pos.min() >= 0 and np.all(pos < np.array([2.0, 1.0, 3.0]))
