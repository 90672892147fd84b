"""The KDK step driver: compute_step + update + fori_loop of the reference.

quantum-stars.py:295-351,454 (stars), quantum-fluid.py:414-459,551 (gas) and
the particle-free step.  ``KDKSimulation.step(n)`` runs n steps on the
device with no host synchronisation; reading ``psi`` (or ``state()``)
applies the owed half kick first, so what a caller sees is always the
reference's end-of-step state.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from . import _lib
from .core import Physics, Plan, ptr


class KDKSimulation:
    """State of one box plus the fused native stepper.

    psi   : complex (nx,ny,nz)
    pos, vel : float64 (np,3) star positions/velocities (stars step)
    mass  : float (common star mass) or float64 (np,) per-particle masses
    gas   : (rho, vx, vy, vz) real fields (gas step)
    v_ext : static external potential added to V in the psi kicks
            (quantum-tidal.py:175-202; psi-only states)
    """

    def __init__(self, plan: Plan, phys: Physics, dt: float, psi: torch.Tensor,
                 pos: torch.Tensor | None = None, vel: torch.Tensor | None = None,
                 mass=0.0, gas=None, t: float = 0.0, v_ext: torch.Tensor | None = None):
        self.plan, self.phys, self.dt = plan, phys, float(dt)
        self._psi = plan.field(psi, "psi", complex_=True).clone()
        dev, rdt = plan.device, plan.real_dtype
        self.t = float(t)
        self.np = 0 if pos is None else int(pos.shape[0])
        if (pos is None) != (vel is None):
            raise ValueError("pos and vel go together")
        if self.np and gas is not None:
            raise ValueError("stars and gas in one simulation are not supported")
        self.pos = self.vel = None
        self.mass_t = None
        self.m_scalar = 0.0
        self._rho_s = None
        self.star_scratch = None
        if self.np:
            self.pos = plan.particles(pos, "pos").clone()
            self.vel = plan.particles(vel, "vel").clone()
            if isinstance(mass, torch.Tensor):
                self.mass_t = plan.particles(mass, "mass", cols=0).clone()
            else:
                self.m_scalar = float(mass)
            self._rho_s = torch.zeros(plan.shape, dtype=rdt, device=dev)
            # tile-binned deposit scratch (counters must start at zero)
            nb = int(_lib.lib().qj_star_scratch_size(plan.handle, self.np))
            self.star_scratch = torch.zeros(nb, dtype=torch.uint8, device=dev)
        self.gas = None
        self._gas_tmp = None
        if gas is not None:
            self.gas = [plan.field(g, n).clone() for g, n in zip(gas, ("rho", "vx", "vy", "vz"))]
            self._gas_tmp = [torch.empty_like(g) for g in self.gas]
        # V (nx,ny,nz) followed by the acceleration grid (nx,ny,nz,3)
        self.grid = torch.empty((1 + _lib.ACC_COMPONENTS) * plan.ncell, dtype=rdt, device=dev)
        self.diag = plan.new_diag()
        self.brho = torch.empty(plan.shape, dtype=rdt, device=dev) if phys.b_coeff else None
        st = _lib.QjState()
        st.psi = self._psi.data_ptr()
        st.rho_s = ptr(self._rho_s).value
        if self.gas is not None:
            for k in range(4):
                st.gas[k] = self.gas[k].data_ptr()
                st.gas_tmp[k] = self._gas_tmp[k].data_ptr()
        st.pos = ptr(self.pos).value
        st.vel = ptr(self.vel).value
        st.mass = ptr(self.mass_t).value
        st.m_scalar = self.m_scalar
        st.np = self.np
        st.grid = self.grid.data_ptr()
        st.brho = ptr(self.brho).value
        st.diag = self.diag.data_ptr()
        st.grid_valid = 0
        st.gas_swapped = 0
        st.pending = 0.0
        st.pending_star = 0.0
        st.star_scratch = ptr(self.star_scratch).value
        st.rho_valid = 0
        self.v_ext = plan.field(v_ext, "v_ext").clone() if v_ext is not None else None
        st.v_ext = ptr(self.v_ext).value
        self._st = st
        self._cfl_out = None
        self.dts = []
        self._params = _lib.QjParams(self.dt, phys.m_per_hbar, phys.four_pi_G, phys.rho_bar,
                                     phys.b_coeff, phys.cs, phys.dc_mode, phys.flux_kind)

    # -- stepping ------------------------------------------------------------
    def set_dt(self, dt: float):
        self.dt = float(dt)
        self._params.dt = self.dt

    def step(self, n: int = 1):
        """Advance n KDK steps (asynchronous on the current stream)."""
        if n <= 0:
            return self
        with torch.cuda.device(self.plan.device):
            _lib.check(_lib.lib().qj_kdk_steps(self.plan.handle, C.byref(self._st),
                                               C.byref(self._params), int(n),
                                               ptr(self.plan.workspace()), self.plan.stream()),
                       "qj_kdk_steps")
        self.t += n * self.dt
        return self

    def step_schedule(self, kick_dts, drift_dts, elapsed: float | None = None):
        """Steps with per-step kick and drift times (e.g. the comoving
        schedule of cosmo.comoving_schedule); psi and gas states.  elapsed,
        when given, is added to t."""
        kick = np.ascontiguousarray(kick_dts, dtype=np.float64)
        drift = np.ascontiguousarray(drift_dts, dtype=np.float64)
        if kick.shape != drift.shape or kick.ndim != 1:
            raise ValueError("kick_dts and drift_dts must be 1-D arrays of one length")
        dp = C.POINTER(C.c_double)
        with torch.cuda.device(self.plan.device):
            _lib.check(_lib.lib().qj_kdk_steps_sched(
                self.plan.handle, C.byref(self._st), C.byref(self._params), int(kick.size),
                kick.ctypes.data_as(dp), drift.ctypes.data_as(dp), ptr(self.plan.workspace()),
                self.plan.stream()), "qj_kdk_steps_sched")
        if elapsed is not None:
            self.t += float(elapsed)
        return self

    def cfl_dt(self, cfl: float = 0.4, dt_max: float | None = None) -> float:
        """CFL time step of the gas state, min(dt_max, cfl dx / max(cs + |v|))
        (quantum-fluid-accretion.py:696-701; isothermal, so c_eff = cs).  The
        maximum is taken on the device over the state the next step starts
        from (owed half kick included); reading it synchronises the host.
        dt_max defaults to self.dt."""
        if self.gas is None:
            raise ValueError("cfl_dt needs a gas state")
        dt_max = self.dt if dt_max is None else float(dt_max)
        if self._cfl_out is None:
            self._cfl_out = torch.empty(2, dtype=torch.float64, device=self.plan.device)
        with torch.cuda.device(self.plan.device):
            _lib.check(_lib.lib().qj_gas_cfl_dt(self.plan.handle, C.byref(self._st),
                                                C.byref(self._params), float(cfl), dt_max,
                                                ptr(self._cfl_out), self.plan.stream()),
                       "qj_gas_cfl_dt")
        return float(self._cfl_out[1].item())

    def step_cfl(self, n: int = 1, cfl: float = 0.4, dt_max: float | None = None):
        """n gas steps, each with the CFL time step of the state it starts
        from (cfl_dt; quantum-fluid-accretion.py:696-701).  The steps taken
        are appended to self.dts."""
        n = max(0, int(n))
        if self.gas is None:
            raise ValueError("step_cfl needs a gas state")
        if n == 0:
            return self
        dt_max = self.dt if dt_max is None else float(dt_max)
        dts = np.zeros(n, dtype=np.float64)
        with torch.cuda.device(self.plan.device):
            rc = _lib.lib().qj_kdk_steps_cfl(self.plan.handle, C.byref(self._st),
                                             C.byref(self._params), n, float(cfl), dt_max,
                                             dts.ctypes.data_as(C.POINTER(C.c_double)),
                                             ptr(self.plan.workspace()), self.plan.stream())
        taken = [float(d) for d in dts if d > 0.0]
        self.dts.extend(taken)
        self.t += math.fsum(taken)
        _lib.check(rc, "qj_kdk_steps_cfl")
        return self

    def finalize(self):
        """Apply the owed half kick (idempotent)."""
        with torch.cuda.device(self.plan.device):
            _lib.check(_lib.lib().qj_kdk_finalize(self.plan.handle, C.byref(self._st),
                                                  C.byref(self._params), self.plan.stream()),
                       "qj_kdk_finalize")
        return self

    # -- state access ------------------------------------------------------
    @property
    def psi(self) -> torch.Tensor:
        self.finalize()
        return self._psi

    def gas_fields(self):
        if self.gas is None:
            return None
        self.finalize()
        return tuple(self._gas_tmp if self._st.gas_swapped else self.gas)

    @property
    def V(self) -> torch.Tensor:
        """Potential of the current state (computed on the first step)."""
        return self.grid[: self.plan.ncell].view(self.plan.shape)

    @property
    def rho_s(self) -> torch.Tensor | None:
        """Star density of the current positions (bin_stars of
        quantum-stars.py:212-247); None without stars.

        Steps with the fused star deposit (_lib.OPT_FUSED_DEPOSIT 2, the
        default) add it to the Poisson source without writing it, so it is
        computed here when read (qj_star_density), into the state's buffer."""
        if self._rho_s is not None and not self._st.rho_valid:
            with torch.cuda.device(self.plan.device):
                _lib.check(_lib.lib().qj_star_density(self.plan.handle, C.byref(self._st),
                                                      self.plan.stream()),
                           "qj_star_density")
        return self._rho_s

    @property
    def a_grid(self) -> torch.Tensor:
        """The reference's a_grid of the current state (stars/gas), (nx,ny,nz,3)
        stacked on the last axis.

        After stars steps with the fused star gather (_lib.OPT_STAR_GATHER,
        the default) the stepper keeps only the accelerations at the stars'
        corners, so the grid is computed here from the state's density
        (|psi|^2 is the same before and after the owed kick) into a new tensor."""
        if self._st.grid_valid == 2:
            from .operators import acceleration_grid
            rho = self._psi.abs().pow(2).to(self.plan.real_dtype) + self.rho_s
            return acceleration_grid(rho, self.plan, self.phys)[1]
        return self.grid[self.plan.ncell:].view(self.plan.shape + (_lib.ACC_COMPONENTS,))

    def a_max(self) -> float:
        return float(self.diag[_lib.DIAG_AMAX].item())

    def dt_s(self) -> float:
        return float(self.diag[_lib.DIAG_DTS].item())

    def nan_seen(self) -> bool:
        return bool(self.diag[_lib.DIAG_NAN].item() != 0.0)

    def state(self) -> dict:
        """The reference's state dict (quantum-stars.py:434-440) after finalize."""
        d = {"psi": self.psi, "t": self.t}
        if self.np:
            d.update(pos=self.pos, vel=self.vel, a_max=self.a_max(), dt_s=self.dt_s())
        if self.gas is not None:
            d.update(zip(("rho", "vx", "vy", "vz"), self.gas_fields()))
        return d


def compute_step(psi, pos, vel, t, a_max, dt_s, *, plan: Plan, phys: Physics, dt: float,
                 m_s=0.0):
    """One stars step with the reference signature (quantum-stars.py:295-330)."""
    sim = KDKSimulation(plan, phys, dt, psi, pos, vel, mass=m_s, t=t)
    sim.diag[_lib.DIAG_AMAX] = float(a_max)
    sim.step(1)
    st = sim.state()
    return st["psi"], st["pos"], st["vel"], st["t"], st["a_max"], st["dt_s"]


def compute_step_gas(psi, rho, vx, vy, vz, t, *, plan: Plan, phys: Physics, dt: float):
    """One gas step with the reference signature (quantum-fluid.py:414-438)."""
    sim = KDKSimulation(plan, phys, dt, psi, gas=(rho, vx, vy, vz), t=t)
    sim.step(1)
    st = sim.state()
    return st["psi"], st["rho"], st["vx"], st["vy"], st["vz"], st["t"]


class BHSimulation(KDKSimulation):
    """Black holes on the field (quantum-bh.py:345-381).

    The holes are particles with per-hole masses m_bh (state), and after every
    step hole 0 accretes from the field at its cell (quantum-bh.py:306-339).
    Accretion changes psi and the masses between a step's closing kick and
    the next step's opening kick, so here each step applies its closing kick,
    accretes, and the next step solves the potential again (two Poisson
    solves per step, as in the reference), instead of merging the kicks.
    """

    def __init__(self, plan: Plan, phys: Physics, dt: float, psi: torch.Tensor,
                 pos: torch.Tensor, vel: torch.Tensor, m_bh: torch.Tensor, G: float, c: float,
                 sigma: float, t: float = 0.0):
        super().__init__(plan, phys, dt, psi, pos, vel, mass=m_bh, t=t)
        self.G, self.c, self.sigma = float(G), float(c), float(sigma)

    def step(self, n: int = 1):
        L = _lib.lib()
        for _ in range(max(0, int(n))):
            super().step(1)
            self.finalize()
            with torch.cuda.device(self.plan.device):
                _lib.check(L.qj_bh_accretion(self.plan.handle, ptr(self._psi), ptr(self.pos),
                                             ptr(self.vel), ptr(self.mass_t), self.np, self.G,
                                             self.c, self.sigma, self.phys.m_per_hbar, self.dt,
                                             self.plan.stream()), "qj_bh_accretion")
            self._st.grid_valid = 0  # the next step solves for the accreted state
        return self

    @property
    def m_bh(self) -> torch.Tensor:
        return self.mass_t

    def state(self) -> dict:
        d = super().state()
        d["m_bh"] = self.mass_t
        return d
