"""KDK steps captured in a graph by the caller (needs a GPU).

A caller may capture qj_kdk_steps on its stream (torch.cuda.graph here):
the step issues no allocation, copy or memset, and the side-stream fork
and join of star steps become graph edges.  A capture records the launches with the host
state of that moment, so it is taken in the steady state (after a first
eager step; gas steps swap their buffers each step, so two steps are
captured for gas).  Each replay, with eager work and host synchronisation
between replays, must equal the same steps taken eagerly by a twin
simulation, bit for bit (deterministic star deposit).
"""

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

C64 = torch.complex64


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda", dtype)


def make(qj, kind, g, side):
    _lib = __import__("importlib").import_module("quantum-jax_amd._lib")
    psi = O.filtered_gaussian_psi(g, 51, sigma_k=8 * np.pi)
    plan = qj.Plan(g.nx, g.ny, g.nz, dtype=C64)
    plan.set_option(_lib.OPT_DETERMINISTIC_DEPOSIT, 1)
    plan.set_option(_lib.OPT_SIDE_STREAM, side)
    if kind == "stars":
        pos, vel, m_s = O.star_ic(g, 3000, 1051, sigma_v=0.2)
        return qj.KDKSimulation(plan, qj.Physics(rho_bar=1.1, four_pi_G=20.0), 1e-3,
                                dev(psi, C64), dev(pos, torch.float64), dev(vel, torch.float64),
                                mass=m_s)
    if kind == "gas":
        rho, vx, vy, vz = O.gas_ic(g, 51, sigma_k=8 * np.pi)
        rb = float(np.mean(np.abs(psi) ** 2 + rho))
        return qj.KDKSimulation(plan, qj.Physics(rho_bar=rb, cs=0.5, flux="hll"), 1e-3,
                                dev(psi, C64), gas=[dev(a, torch.float32) for a in (rho, vx, vy, vz)])
    return qj.KDKSimulation(plan, qj.Physics(four_pi_G=20.0), 1e-3, dev(psi, C64))


def snapshot(sim):
    # the unfinalized buffers: finalizing would apply the owed kick on the
    # host's side of the capture, which the captured steps do not know of
    # (plain steps write V only; the acceleration part of their grid is unused,
    # and after stars steps with the fused star gather a_grid is computed from
    # the state)
    out = [sim._psi.clone(), sim.V.clone()]
    if sim.np or sim.gas is not None:
        out.append(sim.a_grid.clone())
    if sim.np:
        out += [sim.pos.clone(), sim.vel.clone(), sim.rho_s.clone()]
    if sim.gas is not None:
        out += [t.clone() for t in sim.gas + sim._gas_tmp]
    return out


@pytest.mark.parametrize("side", [1, 0])
@pytest.mark.parametrize("kind", ["stars", "gas", "plain"])
def test_captured_steps_replay_like_eager_steps(qj, cuda, kind, side):
    g = O.Grid(32, 32, 32)
    eager, cap = make(qj, kind, g, side), make(qj, kind, g, side)
    per = 2 if kind == "gas" else 1
    eager.step(1)
    cap.step(1)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        cap.step(per)
    torch.cuda.synchronize()
    for rep in range(4):
        graph.replay()
        eager.step(per)
        junk = torch.randn(1 << 16, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        for a, b in zip(snapshot(cap), snapshot(eager)):
            assert torch.equal(a, b), (kind, side, rep)
        assert torch.isfinite(junk)
