"""Device-side diagnostics of a KDK state: total mass and energy.

Neither is computed by the reference (it only asserts time-step bounds,
quantum-stars.py:458-459); they back the mass/energy-drift figures reported
with every benchmark result.  E follows the dynamics of the KDK step:
  i dpsi/dt = [-(1/(2 mph)) lap + mph V] psi,  lap V = 4 pi G (rho - rho_bar),
  dv/dt = -grad V,
  E = vol*sum |grad psi|^2/(2 mph^2) + vol/2*sum V rho + sum m v^2/2.
"""

from __future__ import annotations

import math

import torch

from .core import Physics, Plan
from .operators import bin_particles, get_potential


def _ksq(plan: Plan, dtype):
    def kv(n, L):
        k = torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64, device=plan.device)
        return (2.0 * math.pi / L * k).to(dtype)
    kx = kv(plan.nx, plan.Lx)[:, None, None]
    ky = kv(plan.ny, plan.Ly)[None, :, None]
    kz = kv(plan.nz, plan.Lz)[None, None, :]
    return kx * kx + ky * ky + kz * kz


def mass(plan: Plan, psi, m_p=None, rho_gas=None) -> float:
    M = plan.vol * (psi.abs() ** 2).double().sum().item()
    if m_p is not None:
        M += float(m_p.sum().item()) if isinstance(m_p, torch.Tensor) else float(m_p)
    if rho_gas is not None:
        M += plan.vol * rho_gas.double().sum().item()
    return M


def energy(plan: Plan, phys: Physics, psi, pos=None, vel=None, m_p=None) -> float:
    """Total energy in FP64 (psi is promoted; V is solved on the plan)."""
    psi64 = psi.to(torch.complex128)
    ph = torch.fft.fftn(psi64)
    K = plan.vol * (_ksq(plan, torch.float64) * ph.abs() ** 2).sum().item() / plan.ncell \
        / (2.0 * phys.m_per_hbar ** 2)
    del ph
    rho = (psi.abs() ** 2).to(plan.real_dtype).contiguous()
    if pos is not None and pos.shape[0]:
        rho = rho + bin_particles(pos, m_p, plan)
    p0 = Physics(phys.m_per_hbar, phys.four_pi_G, phys.rho_bar, 0.0, phys.cs, phys.dc_mode)
    V = get_potential(rho.contiguous(), plan, p0)
    W = 0.5 * plan.vol * (V.double() * rho.double()).sum().item()
    E = K + W
    if vel is not None and vel.shape[0]:
        mp = m_p if isinstance(m_p, torch.Tensor) else torch.full(
            (vel.shape[0],), float(m_p), dtype=torch.float64, device=vel.device)
        E += 0.5 * (mp[:, None] * vel * vel).sum().item()
    return E
