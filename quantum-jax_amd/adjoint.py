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

    def __init__(self, plan: Plan, nsteps: int, n_p: int, bins: bool = True):
        dev = plan.device
        N = plan.ncell
        # psi at the start is needed for the first step only (the adjoint merges
        # each closing half kick with the next opening one)
        self.psi0 = torch.empty((1,) + plan.shape, dtype=plan.complex_dtype, device=dev)
        self.psi2 = torch.empty((nsteps,) + plan.shape, dtype=plan.complex_dtype, device=dev)
        G = (1 + _lib.ACC_COMPONENTS) * N  # V, then the acceleration grid
        self.grid1 = torch.empty((nsteps, G), dtype=plan.real_dtype, device=dev)
        self.grid0 = torch.empty(G, dtype=plan.real_dtype, device=dev)
        self.x0 = torch.empty((1, max(n_p, 1), 3), dtype=torch.float64, device=dev)
        self.x1 = torch.empty((nsteps, max(n_p, 1), 3), dtype=torch.float64, device=dev)
        # tile registrations of x1 per step (reused by the adjoint's deposit)
        nb = int(_lib.lib().qj_star_scratch_size(plan.handle, n_p)) if n_p else 0
        self.bins = torch.empty((nsteps, max(nb, 1)), dtype=torch.uint8, device=dev)
        self.steps = (_lib.QjTapeStep * nsteps)()
        for k in range(nsteps):
            t = self.steps[k]
            t.psi0 = self.psi0[0].data_ptr() if k == 0 else None
            t.psi2 = self.psi2[k].data_ptr()
            t.grid1 = self.grid1[k].data_ptr()
            t.x0 = self.x0[0].data_ptr() if n_p and k == 0 else None
            t.x1 = self.x1[k].data_ptr() if n_p else None
            t.bins = self.bins[k].data_ptr() if n_p and bins else None


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
                segment: int, tape_bins: bool):
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
        ctx.tape_bins = tape_bins
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
        tape = _Tape(plan, min(ctx.segment, ctx.nsteps), n_p, ctx.tape_bins)
        starts = list(range(0, ctx.nsteps, ctx.segment))
        for s0, ck in reversed(list(zip(starts, ctx.ckpts))):
            nst = min(ctx.segment, ctx.nsteps - s0)
            _adjoint_segment(plan, phys, dt, ck, mass, nst, tape, adj)
        ctx.ckpts = None
        return psi_bar, x_bar, v_bar, m_bar, None, None, None, None, None, None


def kdk_run(psi, pos, vel, mass, plan: Plan, phys: Physics, dt: float, nsteps: int,
            segment: int = 16, tape_bins: bool = True):
    """nsteps KDK steps (stars), differentiable w.r.t. psi, pos, vel and mass.

    tape_bins: the recompute tapes each step's tile registrations for the
    adjoint's transposed interpolation (False: the adjoint counts them
    again; same gradients up to the order of the additions)."""
    return KDKRun.apply(psi, pos, vel, mass, plan, phys, float(dt), int(nsteps), int(segment),
                        bool(tape_bins))


def density_loss(psi, pos, mass, rho_target, plan: Plan):
    """sum((rho - rho_target)^2) with rho = |psi|^2 + bin_particles(pos, mass),
    the total Poisson source (differentiable)."""
    rho = psi.real ** 2 + psi.imag ** 2 + bin_particles_ad(pos, mass, plan)
    return ((rho - rho_target) ** 2).sum()


class _EnsembleKDKRun(torch.autograd.Function):
    """nsteps KDK steps of many independent members through one Ensemble.

    Inputs are the members' (psi, pos, vel, mass) tensors, flattened as
    psis + poss + vels + masses; outputs are the final psis + poss + vels.
    The backward restores each segment checkpoint into the same member
    buffers (so the recorded ensemble programs repeat across segments) and
    runs the batched tape recording and adjoint (qj_ensemble_record /
    qj_ensemble_adjoint).
    """

    @staticmethod
    def forward(ctx, meta, *flat):
        from .ensemble import Ensemble
        M = meta["n"]
        psis, poss, vels, masses = (flat[k * M:(k + 1) * M] for k in range(4))
        sims = [KDKSimulation(meta["plans"][m], meta["phys"][m], meta["dts"][m], psis[m].detach(),
                              poss[m].detach(), vels[m].detach(), mass=masses[m].detach())
                for m in range(M)]
        ens = Ensemble(sims, mode=meta["mode"])
        nsteps, seg = meta["nsteps"], meta["segment"]
        ckpts = []
        done = 0
        while done < nsteps:
            ens.finalize()  # checkpoints hold end-of-step states (owed kicks applied)
            ckpts.append([(s._psi.clone(), s.pos.clone(), s.vel.clone()) for s in sims])
            n = min(seg, nsteps - done)
            ens.step(n)
            done += n
        ens.finalize()
        ctx.meta, ctx.ens, ctx.sims, ctx.ckpts = meta, ens, sims, ckpts
        ctx.save_for_backward(*[m.detach() for m in masses])
        outs = [s._psi.clone() for s in sims] + [s.pos.clone() for s in sims] + \
               [s.vel.clone() for s in sims]
        return tuple(outs)

    @staticmethod
    def backward(ctx, *grads):
        meta, ens, sims = ctx.meta, ctx.ens, ctx.sims
        M, nsteps, seg = meta["n"], meta["nsteps"], meta["segment"]
        masses = ctx.saved_tensors
        L = _lib.lib()
        adj_t, adj_s = [], (_lib.QjAdjoint * M)()
        for m, s in enumerate(sims):
            plan = s.plan
            gp, gx, gv = grads[m], grads[M + m], grads[2 * M + m]
            pb = (gp if gp is not None else torch.zeros_like(s._psi)).contiguous().clone()
            xb = (gx if gx is not None else torch.zeros_like(s.pos)).contiguous().clone()
            vb = (gv if gv is not None else torch.zeros_like(s.vel)).contiguous().clone()
            mb = torch.zeros_like(masses[m])
            sc = torch.empty(int(L.qj_adjoint_scratch_size(plan.handle)), dtype=torch.uint8,
                             device=plan.device)
            adj_t.append((pb, xb, vb, mb, sc))
            adj_s[m] = _lib.QjAdjoint(pb.data_ptr(), xb.data_ptr(), vb.data_ptr(), mb.data_ptr(),
                                      sc.data_ptr())
        tapes = [_Tape(s.plan, min(seg, nsteps), s.np) for s in sims]
        tape_ptrs = (C.POINTER(_lib.QjTapeStep) * M)(
            *[C.cast(t.steps, C.POINTER(_lib.QjTapeStep)) for t in tapes])
        grid0s = (_lib.vp * M)(*[t.grid0.data_ptr() for t in tapes])
        starts = list(range(0, nsteps, seg))
        with torch.cuda.device(ens.device):
            stream = C.c_void_p(torch.cuda.current_stream(ens.device).cuda_stream)
            for s0, ck in reversed(list(zip(starts, ctx.ckpts))):
                nst = min(seg, nsteps - s0)
                for s, (p, x, v) in zip(sims, ck):
                    s._psi.copy_(p)
                    s.pos.copy_(x)
                    s.vel.copy_(v)
                    st = s._st
                    st.grid_valid, st.gas_swapped, st.pending, st.pending_star = 0, 0, 0.0, 0.0
                _lib.check(L.qj_ensemble_record(ens._h, nst, tape_ptrs, grid0s, stream),
                           "qj_ensemble_record")
                _lib.check(L.qj_ensemble_adjoint(ens._h, nst, tape_ptrs, grid0s, adj_s, stream),
                           "qj_ensemble_adjoint")
        ctx.ckpts = None
        ens.close()
        res = [None] + [a[0] for a in adj_t] + [a[1] for a in adj_t] + [a[2] for a in adj_t] + \
              [a[3] for a in adj_t]
        return tuple(res)


def ensemble_kdk_run(psis, poss, vels, masses, plans, physs, dts, nsteps: int,
                     segment: int = 16, mode: str = "graph"):
    """Differentiable nsteps KDK steps of M independent boxes, stepped as one
    Ensemble (batched per shape, graph-replayed) forward and backward.

    psis, poss, vels, masses : per-member tensors (masses float64 (np,));
    plans : one Plan per member; physs, dts : per-member Physics and dt.
    Returns (psis_out, poss_out, vels_out), lists of tensors.
    """
    M = len(psis)
    meta = {"n": M, "plans": list(plans), "phys": list(physs), "dts": [float(d) for d in dts],
            "nsteps": int(nsteps), "segment": int(segment), "mode": mode}
    out = _EnsembleKDKRun.apply(meta, *psis, *poss, *vels, *masses)
    return list(out[:M]), list(out[M:2 * M]), list(out[2 * M:])
