"""Kick/drift variants of the reference: comoving cosmology and a static
external potential.

Comoving step (quantum-cosmo.py:149-197): between scale factors a and a+da
the kicks use dt * pot_fac / 2 and the drift dt * kin_fac, with
kin_fac = (a^-2 + (a+da)^-2)/2, pot_fac = (a^-1 + (a+da)^-1)/2 and dt from a
1000-point midpoint quadrature of 1/adot.  ``comoving_schedule`` turns a run
of steps into the per-step kick and drift times of
``KDKSimulation.step_schedule`` (qj_kdk_steps_sched), which keeps the merged
kicks: the closing kick of step k and the opening kick of step k+1 share the
potential.

External potential (quantum-tidal.py:175-202): a static V_ext added to V in
the psi kicks (``KDKSimulation(..., v_ext=...)``); ``point_mass_potential``
is the reference's softened point mass.
"""

from __future__ import annotations

import numpy as np
import torch

from .core import Plan


def compute_dt(a: float, da: float, H0: float, omega_m: float, omega_l: float,
               n_quad: int = 1000) -> float:
    """quantum-cosmo.py:149-163 (host, FP64)."""
    h = 1.0 / n_quad
    lin = np.linspace(0.5 * h, 1.0 - 0.5 * h, n_quad)

    def t_of(aa):
        at = aa * lin
        return float(np.mean(1.0 / (H0 * np.sqrt(omega_m / at + omega_l * at ** 2)))) * aa

    return t_of(a + da) - t_of(a)


def comoving_schedule(a_start: float, da: float, nsteps: int, H0: float, omega_m: float,
                      omega_l: float):
    """Per-step (kick_dt, drift_dt) = (dt pot_fac, dt kin_fac) of nsteps
    comoving steps from a_start, and the final scale factor."""
    kick = np.empty(nsteps)
    drift = np.empty(nsteps)
    a = float(a_start)
    for k in range(nsteps):
        kin_fac = 0.5 * (a ** -2 + (a + da) ** -2)
        pot_fac = 0.5 * (a ** -1 + (a + da) ** -1)
        dt = compute_dt(a, da, H0, omega_m, omega_l)
        kick[k] = dt * pot_fac
        drift[k] = dt * kin_fac
        a += da
    return kick, drift, a


def point_mass_potential(plan: Plan, M: float, G: float) -> torch.Tensor:
    """-G M / (r + dx/2) with r from the box centre (quantum-tidal.py:175-179),
    on the plan's device in its real type (built once, in FP64)."""
    x = [np.linspace(0.5 * L / n, L - 0.5 * L / n, n)
         for n, L in ((plan.nx, plan.Lx), (plan.ny, plan.Ly), (plan.nz, plan.Lz))]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    r = np.sqrt((X - 0.5 * plan.Lx) ** 2 + (Y - 0.5 * plan.Lx) ** 2 + (Z - 0.5 * plan.Lx) ** 2)
    V = -G * M / (r + 0.5 * plan.dx)
    return torch.as_tensor(V).to(plan.device, plan.real_dtype).contiguous()
