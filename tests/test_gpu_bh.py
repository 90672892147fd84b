"""Black-hole accretion and the black-hole step against the oracle (needs a GPU)."""

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda", dtype)


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def case(seed=8):
    g = O.Grid(16, 16, 16)
    psi = O.filtered_gaussian_psi(g, seed, sigma_k=6 * np.pi)
    pos = np.array([[0.31, 0.52, 0.77], [0.1, 0.2, 0.3], [0.9, 0.05, 0.5]])
    vel = np.array([[0.3, -0.2, 0.1], [0.0, 0.1, 0.0], [-0.2, 0.0, 0.3]])
    m_bh = np.array([0.05, 0.02, 0.03])
    return g, psi, pos, vel, m_bh


@pytest.mark.parametrize("dtype", [C64, C128])
def test_accretion_matches_oracle(qj, cuda, dtype):
    g, psi, pos, vel, m_bh = case()
    kw = dict(G=1.0, c=50.0, sigma=0.4)
    p_ref, m_ref = O.do_accretion(psi, pos, vel, m_bh, g, O.Physics(m_per_hbar=2.0),
                                  dt=1e-3, **kw)
    plan = qj.Plan(16, dtype=dtype)
    p, m = qj.do_accretion(dev(psi, dtype), dev(pos, torch.float64), dev(vel, torch.float64),
                           dev(m_bh, torch.float64), 1e-3, plan, qj.Physics(m_per_hbar=2.0), **kw)
    assert rel(p, p_ref) < (1e-7 if dtype == C64 else 1e-15)
    assert np.allclose(m.cpu().numpy(), m_ref, rtol=1e-13 if dtype == C128 else 1e-6, atol=0)


def test_bh_steps_match_oracle(qj, cuda):
    g, psi, pos, vel, m_bh = case(9)
    ph = O.Physics(m_per_hbar=2.0, four_pi_G=20.0, rho_bar=1.0)
    kw = dict(G=1.0, c=50.0, sigma=0.4)
    dt, nsteps = 1e-3, 4
    s = (psi, m_bh, pos, vel, 0.0)
    for _ in range(nsteps):
        s = O.compute_step_bh(*s, dt, g, ph, **kw)
    plan = qj.Plan(16, dtype=C128)
    sim = qj.BHSimulation(plan, qj.Physics(m_per_hbar=2.0, four_pi_G=20.0, rho_bar=1.0), dt,
                          dev(psi, C128), dev(pos, torch.float64), dev(vel, torch.float64),
                          dev(m_bh, torch.float64), **kw)
    sim.step(nsteps)
    st = sim.state()
    assert rel(st["psi"], s[0]) < 1e-11
    assert rel(st["m_bh"], s[1]) < 1e-12
    assert rel(st["pos"], s[2]) < 1e-12
    assert rel(st["vel"], s[3]) < 1e-10
    assert st["t"] == pytest.approx(s[4])
