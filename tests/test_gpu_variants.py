"""Kick/drift variants against the FP64 oracle (needs a GPU).

quantum-tidal.py: a static external potential (softened point mass) added
to V in the psi kicks; quantum-cosmo.py: the comoving step, whose kicks use
dt pot_fac and drift dt kin_fac with dt from the quadrature of 1/adot, run
through the per-step schedule entry point (merged kicks).
"""

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128
TOL = {C64: 2e-5, C128: 1e-11}


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda", dtype)


@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("n", [16, 32])
def test_external_potential_steps_match_oracle(qj, cuda, dtype, n):
    g = O.Grid(n, n, n)
    ph = O.Physics(four_pi_G=10.0, m_per_hbar=1.0, rho_bar=1.0)
    psi = O.filtered_gaussian_psi(g, 9, sigma_k=6 * np.pi)
    M, G, dt, nsteps = 3.0, 0.8, 1e-3, 5
    Vx = O.external_point_mass(g, M, G)
    p = psi
    for _ in range(nsteps):
        p = O.compute_step_tidal(p, dt, Vx, g, ph)
    plan = qj.Plan(n, dtype=dtype)
    v_ext = qj.cosmo.point_mass_potential(plan, M, G)
    assert rel(v_ext, Vx) < (1e-7 if dtype == C64 else 1e-15)
    sim = qj.KDKSimulation(plan, qj.Physics(four_pi_G=10.0, rho_bar=1.0), dt,
                           dev(psi, dtype), v_ext=v_ext)
    sim.step(nsteps)
    assert rel(sim.psi, p) < TOL[dtype]


@pytest.mark.parametrize("dtype", [C64, C128])
def test_comoving_schedule_matches_oracle(qj, cuda, dtype):
    g = O.Grid(32, 32, 32)
    ph = O.Physics(four_pi_G=10.0, m_per_hbar=1.0, rho_bar=1.0)
    psi = O.filtered_gaussian_psi(g, 10, sigma_k=6 * np.pi)
    a0, da, nsteps, H0, om, ol = 0.1, 2e-3, 6, 0.5, 0.3, 0.7
    p, a = psi, a0
    for _ in range(nsteps):
        p, a = O.compute_step_cosmo(p, a, da, g, ph, H0, om, ol)
    kick, drift, a_end = qj.cosmo.comoving_schedule(a0, da, nsteps, H0, om, ol)
    assert a_end == pytest.approx(a)
    plan = qj.Plan(32, dtype=dtype)
    sim = qj.KDKSimulation(plan, qj.Physics(four_pi_G=10.0, rho_bar=1.0), 0.0, dev(psi, dtype))
    sim.step_schedule(kick[:2], drift[:2])
    sim.step_schedule(kick[2:], drift[2:])  # the owed kick carries across calls
    assert rel(sim.psi, p) < TOL[dtype]


def test_variants_reject_stars(qj, cuda):
    g = O.Grid(16, 16, 16)
    plan = qj.Plan(16)
    psi = dev(O.filtered_gaussian_psi(g, 1), C64)
    pos, vel, m = O.star_ic(g, 10, 2)
    sim = qj.KDKSimulation(plan, qj.Physics(), 1e-3, psi, dev(pos, torch.float64),
                           dev(vel, torch.float64), mass=m,
                           v_ext=torch.zeros(g.shape, dtype=torch.float32, device="cuda"))
    with pytest.raises(qj.QjError):
        sim.step(1)
    sim2 = qj.KDKSimulation(qj.Plan(16), qj.Physics(), 1e-3, psi, dev(pos, torch.float64),
                            dev(vel, torch.float64), mass=m)
    with pytest.raises(qj.QjError):
        sim2.step_schedule([1e-3], [1e-3])
