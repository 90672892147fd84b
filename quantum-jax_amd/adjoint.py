"""Hand-written adjoint of the KDK step behind torch.autograd.

The reference calls its solver differentiable (README.md:7) but never takes a
gradient; here the whole split-step is differentiated by hand on the device:
conjugate phase kicks, the transposed drift (the drift of -dt), the Poisson
operator (symmetric) applied to the adjoint source V_bar + div a_bar, the
transposes of the CIC interpolation and deposit, and the leapfrog.

``kdk_run`` advances ``nsteps`` steps and is differentiable w.r.t. the
initial psi (complex), star positions, velocities and per-star masses.  The
forward keeps only segment checkpoints (every ``segment`` steps); the
backward recomputes each segment while recording the per-step tape, then
runs the adjoint steps over it.

Complex gradients follow PyTorch's convention, grad = dL/dRe + i dL/dIm
(JAX's jax.grad returns the complex conjugate of this).
"""

from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from .core import Physics, Plan, ptr
from .simulation import KDKSimulation


class BinParticles(torch.autograd.Function):
    """rho = bin_particles(pos, m) with gradients w.r.t. pos and m.

    quantum-bh.py:223-258; backward = the transpose of the deposit.
    """

    @staticmethod
    def forward(ctx, pos, mass, plan: Plan):
        from .operators import bin_particles
        rho = bin_particles(pos.contiguous(), mass.contiguous(), plan)
        ctx.save_for_backward(pos, mass)
        ctx.plan = plan
        return rho

    @staticmethod
    def backward(ctx, g_rho):
        pos, mass = ctx.saved_tensors
        plan = ctx.plan
        g = g_rho.contiguous().to(plan.real_dtype)
        m_bar = torch.zeros_like(mass)
        x_bar = torch.zeros_like(pos)
        with torch.cuda.device(plan.device):
            _lib.check(_lib.lib().qj_deposit_adjoint(
                plan.handle, ptr(pos), ptr(mass), 0.0, pos.shape[0], ptr(g), ptr(m_bar),
                ptr(x_bar), plan.stream()), "qj_deposit_adjoint")
        return x_bar, m_bar, None


def bin_particles_ad(pos, mass, plan: Plan):
    return BinParticles.apply(pos, mass, plan)


class _Tape:
    """Device buffers of one recorded segment (reused across segments)."""

    def __init__(self, plan: Plan, nsteps: int, n_p: int):
        dev = plan.device
        N = plan.ncell
        self.psi0 = torch.empty((nsteps,) + plan.shape, dtype=plan.complex_dtype, device=dev)
        self.psi2 = torch.empty_like(self.psi0)
        self.grid1 = torch.empty((nsteps, 5 * N), dtype=plan.real_dtype, device=dev)
        self.grid0 = torch.empty(5 * N, dtype=plan.real_dtype, device=dev)
        self.x0 = torch.empty((nsteps, max(n_p, 1), 3), dtype=torch.float64, device=dev)
        self.x1 = torch.empty_like(self.x0)
        self.steps = (_lib.QjTapeStep * nsteps)()
        for k in range(nsteps):
            t = self.steps[k]
            t.psi0 = self.psi0[k].data_ptr()
            t.psi2 = self.psi2[k].data_ptr()
            t.grid1 = self.grid1[k].data_ptr()
            t.x0 = self.x0[k].data_ptr() if n_p else None
            t.x1 = self.x1[k].data_ptr() if n_p else None


def _adjoint_segment(plan: Plan, phys: Physics, dt: float, ck, mass, nst: int, tape: _Tape,
                     adj: _lib.QjAdjoint):
    psi, pos, vel = ck
    sim = KDKSimulation(plan, phys, dt, psi, pos, vel, mass=mass)
    L = _lib.lib()
    with torch.cuda.device(plan.device):
        _lib.check(L.qj_kdk_record(plan.handle, C.byref(sim._st), C.byref(sim._params), nst,
                                   tape.steps, ptr(tape.grid0), ptr(plan.workspace()),
                                   plan.stream()), "qj_kdk_record")
        _lib.check(L.qj_kdk_adjoint(plan.handle, C.byref(sim._st), C.byref(sim._params), nst,
                                    tape.steps, ptr(tape.grid0), C.byref(adj),
                                    ptr(plan.workspace()), plan.stream()), "qj_kdk_adjoint")


class KDKRun(torch.autograd.Function):
    @staticmethod
    def forward(ctx, psi, pos, vel, mass, plan: Plan, phys: Physics, dt: float, nsteps: int,
                segment: int):
        sim = KDKSimulation(plan, phys, dt, psi.detach(), pos.detach(), vel.detach(),
                            mass=mass.detach())
        ckpts = []
        done = 0
        while done < nsteps:
            ckpts.append((sim.psi.clone(), sim.pos.clone(), sim.vel.clone()))
            n = min(segment, nsteps - done)
            sim.step(n)
            done += n
        st = sim.state()
        ctx.plan, ctx.phys, ctx.dt, ctx.nsteps, ctx.segment = plan, phys, dt, nsteps, segment
        ctx.ckpts = ckpts
        ctx.save_for_backward(mass.detach())
        return st["psi"].clone(), st["pos"].clone(), st["vel"].clone()

    @staticmethod
    def backward(ctx, g_psi, g_pos, g_vel):
        (mass,) = ctx.saved_tensors
        plan, phys, dt = ctx.plan, ctx.phys, ctx.dt
        n_p = mass.shape[0]
        dev = plan.device
        psi_bar = (g_psi if g_psi is not None else
                   torch.zeros(plan.shape, dtype=plan.complex_dtype, device=dev)).contiguous().clone()
        x_bar = (g_pos if g_pos is not None else
                 torch.zeros((n_p, 3), dtype=torch.float64, device=dev)).contiguous().clone()
        v_bar = (g_vel if g_vel is not None else
                 torch.zeros((n_p, 3), dtype=torch.float64, device=dev)).contiguous().clone()
        m_bar = torch.zeros(n_p, dtype=torch.float64, device=dev)
        scratch = torch.empty(int(_lib.lib().qj_adjoint_scratch_size(plan.handle)),
                              dtype=torch.uint8, device=dev)
        adj = _lib.QjAdjoint(psi_bar.data_ptr(), x_bar.data_ptr(), v_bar.data_ptr(),
                             m_bar.data_ptr(), scratch.data_ptr())
        tape = _Tape(plan, min(ctx.segment, ctx.nsteps), n_p)
        starts = list(range(0, ctx.nsteps, ctx.segment))
        for s0, ck in reversed(list(zip(starts, ctx.ckpts))):
            nst = min(ctx.segment, ctx.nsteps - s0)
            _adjoint_segment(plan, phys, dt, ck, mass, nst, tape, adj)
        ctx.ckpts = None
        return psi_bar, x_bar, v_bar, m_bar, None, None, None, None, None


def kdk_run(psi, pos, vel, mass, plan: Plan, phys: Physics, dt: float, nsteps: int,
            segment: int = 16):
    """nsteps KDK steps (stars), differentiable w.r.t. psi, pos, vel and mass."""
    return KDKRun.apply(psi, pos, vel, mass, plan, phys, float(dt), int(nsteps), int(segment))


def density_loss(psi, pos, mass, rho_target, plan: Plan):
    """sum((rho - rho_target)^2) with rho = |psi|^2 + bin_particles(pos, mass),
    the total Poisson source (differentiable)."""
    rho = psi.real ** 2 + psi.imag ** 2 + bin_particles_ad(pos, mass, plan)
    return ((rho - rho_target) ** 2).sum()
