"""The reference operator surface, one function per reference operator.

Each function validates its tensors, calls the gfx950 library through the
C ABI on torch's current stream and returns new tensors (the reference's
arrays are immutable, so inputs are never modified unless a name ends in
an underscore).  Module globals of the reference are the ``plan`` and
``phys`` arguments.
"""

from __future__ import annotations

import torch

from . import _lib
from .core import Physics, Plan, ptr


def _L():
    return _lib.lib()


def get_potential(rho: torch.Tensor, plan: Plan, phys: Physics) -> torch.Tensor:
    """V = Re IFFT(-FFT(4 pi G (rho - rho_bar)) / (k^2 + [k^2 == 0])) [+ b rho].

    quantum-stars.py:189-193; quantum-fluid.py:171-175.
    """
    plan.field(rho, "rho")
    V = torch.empty_like(rho)
    with torch.cuda.device(plan.device):
        _lib.check(_L().qj_get_potential(plan.handle, ptr(rho), phys.rho_bar, phys.four_pi_G,
                                         phys.b_coeff, phys.dc_mode, ptr(V),
                                         ptr(plan.workspace()), plan.stream()),
                   "get_potential")
    return V


def _new_a_grid(plan: Plan) -> torch.Tensor:
    return torch.empty(plan.shape + (_lib.ACC_COMPONENTS,), dtype=plan.real_dtype,
                       device=plan.device)


def acceleration_grid(rho: torch.Tensor, plan: Plan, phys: Physics, diag=None):
    """(V, a_grid): a_grid[..., d] = -Re IFFT(1j k_d V_hat), d = x, y, z.

    quantum-stars.py:254-259 (the reference's a_grid, stacked on the last
    axis).  max |a| is folded into diag[0] when diag is given.
    """
    plan.field(rho, "rho")
    V = torch.empty_like(rho)
    a = _new_a_grid(plan)
    with torch.cuda.device(plan.device):
        _lib.check(_L().qj_acceleration_grid(plan.handle, ptr(rho), phys.rho_bar,
                                             phys.four_pi_G, phys.dc_mode, ptr(V), ptr(a),
                                             ptr(diag), ptr(plan.workspace()), plan.stream()),
                   "acceleration_grid")
    return V, a


def kick(psi: torch.Tensor, V: torch.Tensor, coef: float, plan: Plan) -> torch.Tensor:
    """psi * exp(-1j coef V), coef = m_per_hbar*dt/2; quantum-stars.py:301."""
    out = plan.field(psi, "psi", complex_=True).clone()
    plan.field(V, "V")
    with torch.cuda.device(plan.device):
        _lib.check(_L().qj_kick(plan.handle, ptr(out), ptr(V), float(coef), plan.stream()),
                   "kick")
    return out


def drift(psi: torch.Tensor, dt: float, plan: Plan, phys: Physics) -> torch.Tensor:
    """IFFT(exp(-1j dt k^2/(2 m_per_hbar)) FFT psi); quantum-stars.py:307-309."""
    out = plan.field(psi, "psi", complex_=True).clone()
    with torch.cuda.device(plan.device):
        _lib.check(_L().qj_drift(plan.handle, ptr(out), float(dt), phys.m_per_hbar,
                                 plan.stream()), "drift")
    return out


def get_cic_indices_and_weights(pos: torch.Tensor, plan: Plan):
    """(i, ip1, w_i, w_ip1); quantum-stars.py:200-209 (int32 indices as the
    reference's default-precision ``astype(int)``)."""
    plan.particles(pos, "pos")
    n = pos.shape[0]
    i = torch.empty((n, 3), dtype=torch.int32, device=plan.device)
    ip1 = torch.empty_like(i)
    w_i = torch.empty((n, 3), dtype=torch.float64, device=plan.device)
    w_ip1 = torch.empty_like(w_i)
    with torch.cuda.device(plan.device):
        _lib.check(_L().qj_cic_indices(plan.handle, ptr(pos), n, ptr(i), ptr(ip1), ptr(w_i),
                                       ptr(w_ip1), plan.stream()), "cic_indices")
    return i, ip1, w_i, w_ip1


def bin_particles(pos: torch.Tensor, m, plan: Plan, out: torch.Tensor | None = None,
                  accumulate: bool = False) -> torch.Tensor:
    """CIC deposit with per-particle (tensor) or common (float) mass.

    quantum-bh.py:223-258; quantum-stars.py:212-247 for a scalar m.
    """
    plan.particles(pos, "pos")
    n = pos.shape[0]
    mass = None
    m_scalar = 0.0
    if isinstance(m, torch.Tensor):
        mass = plan.particles(m, "m", cols=0)
        if mass.shape[0] != n:
            raise ValueError("m must have one entry per particle")
    else:
        m_scalar = float(m)
    if out is None:
        out = torch.empty(plan.shape, dtype=plan.real_dtype, device=plan.device)
        accumulate = False
    else:
        plan.field(out, "out")
    with torch.cuda.device(plan.device):
        _lib.check(_L().qj_bin_particles(plan.handle, ptr(pos), ptr(mass), m_scalar, n,
                                         ptr(out), int(bool(accumulate)), plan.stream()),
                   "bin_particles")
    return out


def bin_stars(pos: torch.Tensor, m_s: float, plan: Plan) -> torch.Tensor:
    """quantum-stars.py:212-247."""
    return bin_particles(pos, float(m_s), plan)


def interp(pos: torch.Tensor, a_grid: torch.Tensor, plan: Plan) -> torch.Tensor:
    """8-corner CIC gather of a_grid ((nx,ny,nz,3)); quantum-stars.py:262-287."""
    plan.particles(pos, "pos")
    shape = plan.shape + (_lib.ACC_COMPONENTS,)
    if (a_grid.shape != shape or a_grid.dtype != plan.real_dtype
            or a_grid.device != plan.device or not a_grid.is_contiguous()):
        raise ValueError(f"a_grid must be a contiguous {plan.real_dtype} tensor of shape "
                         f"{shape} on {plan.device}")
    acc = torch.empty((pos.shape[0], 3), dtype=torch.float64, device=plan.device)
    with torch.cuda.device(plan.device):
        _lib.check(_L().qj_interp(plan.handle, ptr(pos), pos.shape[0], ptr(a_grid), ptr(acc),
                                  plan.stream()), "interp")
    return acc


def get_acceleration(pos: torch.Tensor, rho: torch.Tensor, plan: Plan, phys: Physics):
    """(acc, a_max) of quantum-stars.py:250-288.  a_max is a device scalar."""
    plan.particles(pos, "pos")
    plan.field(rho, "rho")
    diag = plan.new_diag()
    a = _new_a_grid(plan)
    acc = torch.empty((pos.shape[0], 3), dtype=torch.float64, device=plan.device)
    with torch.cuda.device(plan.device):
        _lib.check(_L().qj_get_acceleration(plan.handle, ptr(pos), pos.shape[0], ptr(rho),
                                            phys.rho_bar, phys.four_pi_G, phys.dc_mode,
                                            ptr(acc), ptr(a), ptr(diag),
                                            ptr(plan.workspace()), plan.stream()),
                   "get_acceleration")
    return acc, diag[_lib.DIAG_AMAX]


def apply_grav_accel(rho, vx, vy, vz, dt: float, plan: Plan, phys: Physics):
    """quantum-fluid.py:267-278 -> (vx, vy, vz) (new tensors)."""
    plan.field(rho, "rho")
    out = [plan.field(v, n).clone() for v, n in ((vx, "vx"), (vy, "vy"), (vz, "vz"))]
    a = _new_a_grid(plan)
    with torch.cuda.device(plan.device):
        _lib.check(_L().qj_apply_grav_accel(plan.handle, ptr(rho), ptr(out[0]), ptr(out[1]),
                                            ptr(out[2]), float(dt), phys.rho_bar,
                                            phys.four_pi_G, phys.dc_mode, ptr(a), None,
                                            ptr(plan.workspace()), plan.stream()),
                   "apply_grav_accel")
    return tuple(out)


def solve_hydro(rho, vx, vy, vz, dt: float, plan: Plan, phys: Physics):
    """MUSCL-Hancock isothermal update; quantum-fluid.py:281-349 (face flux
    phys.flux: the reference's Rusanov form, or HLL)."""
    ins = [plan.field(v, n) for v, n in ((rho, "rho"), (vx, "vx"), (vy, "vy"), (vz, "vz"))]
    outs = [torch.empty_like(v) for v in ins]
    with torch.cuda.device(plan.device):
        _lib.check(_L().qj_solve_hydro(plan.handle, *[ptr(v) for v in ins], float(dt),
                                       phys.cs, phys.flux_kind, *[ptr(v) for v in outs],
                                       plan.stream()),
                   "solve_hydro")
    return tuple(outs)


__all__ = [
    "get_potential", "acceleration_grid", "kick", "drift", "get_cic_indices_and_weights",
    "bin_particles", "bin_stars", "interp", "get_acceleration", "apply_grav_accel",
    "solve_hydro",
]


def do_accretion(psi, pos, vel, m_bh, dt: float, plan: Plan, phys: Physics, G: float, c: float,
                 sigma: float):
    """(psi, m_bh) after accretion onto black hole 0; quantum-bh.py:306-339."""
    psi = plan.field(psi, "psi", complex_=True).clone()
    plan.particles(pos, "pos")
    plan.particles(vel, "vel")
    m = plan.particles(m_bh, "m_bh", cols=0).clone()
    with torch.cuda.device(plan.device):
        _lib.check(_lib.lib().qj_bh_accretion(plan.handle, ptr(psi), ptr(pos), ptr(vel), ptr(m),
                                              m.shape[0], float(G), float(c), float(sigma),
                                              phys.m_per_hbar, float(dt), plan.stream()),
                   "qj_bh_accretion")
    return psi, m
