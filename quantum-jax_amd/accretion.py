"""Black-hole operators of the gas-accretion variant (quantum-fluid-accretion.py).

  * ``accel_lp`` / ``bh_accel_lp``: the low-pass-filtered gravity that kicks
    the hole, accel_from_Vhat(get_potential_bg(rho, rho_bg)[0] * LP) with
    LP = exp(-(k^2 / (k_cut^2 + 1e-30))^4) (:140-146,182-191,719-725), sampled
    at the hole by CIC (:419-438);
  * ``bh_bondi_step_ambient``: the fixed-grid Bondi sink with an ambient
    shell (:440-510), on the device in one workgroup;
  * ``seed_mass_from_time``: the seed-mass ramp (:682-690), host scalar.

The potential with a per-step background (get_potential_bg) is
``operators.get_potential`` with ``Physics(rho_bar=rho_bg)``.
"""

from __future__ import annotations

import math

import torch

from . import _lib
from .core import Physics, Plan, ptr
from .operators import _new_a_grid, interp

LAMBDA_ISO = math.exp(1.5) / 4.0  # quantum-fluid-accretion.py:63


def accel_lp(rho: torch.Tensor, plan: Plan, phys: Physics, k_cut: float, rho_bg=None):
    """(V_lp, a_grid) of the filtered potential; rho_bg defaults to phys.rho_bar."""
    plan.field(rho, "rho")
    rb = phys.rho_bar if rho_bg is None else float(rho_bg)
    V = torch.empty_like(rho)
    a = _new_a_grid(plan)
    with torch.cuda.device(plan.device):
        _lib.check(_lib.lib().qj_acceleration_grid_lp(plan.handle, ptr(rho), rb, phys.four_pi_G,
                                                      float(k_cut), ptr(V), ptr(a),
                                                      ptr(plan.workspace()), plan.stream()),
                   "qj_acceleration_grid_lp")
    return V, a


def bh_accel_lp(rho: torch.Tensor, x_bh: torch.Tensor, plan: Plan, phys: Physics, k_cut: float,
                rho_bg=None) -> torch.Tensor:
    """The filtered acceleration at the hole(s): x_bh float64 (n, 3) -> (n, 3)."""
    _, a = accel_lp(rho, plan, phys, k_cut, rho_bg)
    return interp(x_bh, a, plan)


def bh_bondi_step_ambient(rho: torch.Tensor, bh: torch.Tensor, plan: Plan, cs: float, G: float,
                          dt: float, r_acc_mult: float = 2.5, lam: float = LAMBDA_ISO,
                          fmax=None, ann_lo: float = 1.1, ann_hi: float = 1.8, v=None):
    """quantum-fluid-accretion.py:440-510 (defaults of :62-66).

    rho: gas density (not modified); bh: float64 (7,) device tensor (M_eff,
    x, y, z, vx, vy, vz); v: optional (vx, vy, vz) gas velocity fields for
    the Bondi-Hoyle-Lyttleton rate.  Returns (rho_new, info) with info a
    float64 (8,) device tensor: dMdt_target, rho_inf, r_B, dM_cap, r_acc,
    M_avail_kernel, dM_uncl, shell cell count.
    """
    plan.field(rho, "rho")
    if bh.dtype != torch.float64 or bh.device != plan.device or bh.numel() != 7:
        raise ValueError("bh must be a float64 tensor of 7 values on the plan's device")
    bh = bh.contiguous()
    vs = [None, None, None]
    if v is not None:
        vs = [plan.field(t, n) for t, n in zip(v, ("vx", "vy", "vz"))]
    out = rho.clone()
    info = torch.empty(8, dtype=torch.float64, device=plan.device)
    with torch.cuda.device(plan.device):
        _lib.check(_lib.lib().qj_bondi_sink(plan.handle, ptr(out), ptr(vs[0]), ptr(vs[1]),
                                            ptr(vs[2]), ptr(bh), float(cs), float(G), float(dt),
                                            float(r_acc_mult), float(lam),
                                            -1.0 if fmax is None else float(fmax),
                                            float(ann_lo), float(ann_hi), ptr(info),
                                            plan.stream()), "qj_bondi_sink")
    return out, info


def seed_mass_from_time(t: float, placed: float, init_m: float, inject_t: float,
                        ramp_center=None, ramp_w=None) -> float:
    """quantum-fluid-accretion.py:682-690: init_m s(t) for a placed hole; s is
    the tanh ramp (1 + tanh((t - ramp_center) / (ramp_w + 1e-30))) / 2 from
    inject_t on, or the step [t >= inject_t] when ramp_center is None."""
    on = 1.0 if placed > 0.5 else 0.0
    if ramp_center is not None:
        s = 0.5 * (1.0 + math.tanh((t - ramp_center) / (ramp_w + 1e-30)))
        return init_m * s * (1.0 if t >= inject_t else 0.0) * on
    return init_m * (1.0 if t >= inject_t else 0.0) * on


__all__ = ["accel_lp", "bh_accel_lp", "bh_bondi_step_ambient", "seed_mass_from_time",
           "LAMBDA_ISO"]
