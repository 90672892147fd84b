"""The deterministic star deposit (QJ_OPT_DETERMINISTIC_DEPOSIT, needs a GPU).

The tile-binned deposit normally adds corner values with LDS float atomics,
so two stars runs differ by rounding.  With the option on, each tile adds
64-bit fixed-point integers, whose sum is independent of the order: stars
runs must then reproduce bit for bit (the rest of the step is already
order-independent: fixed-order FP64 source sums, max/min diagnostics), the
deposited density must match the oracle's bin_particles, and an ensemble
of star boxes must match member-by-member stepping bit for bit.
"""

import importlib

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda", dtype)


def rel(a, b):
    return ((a - b).norm() / b.norm()).item()


def _run(qj, g, dtype, psi, pos, vel, m, det, nsteps=3):
    _lib = importlib.import_module("quantum-jax_amd._lib")
    plan = qj.Plan(g.nx, g.ny, g.nz, L=(g.Lx, g.Ly, g.Lz), dtype=dtype)
    plan.set_option(_lib.OPT_DETERMINISTIC_DEPOSIT, det)
    assert plan.get_option(_lib.OPT_DETERMINISTIC_DEPOSIT) == det
    sim = qj.KDKSimulation(plan, qj.Physics(rho_bar=1.1, four_pi_G=20.0), 1e-3, dev(psi, dtype),
                           dev(pos, torch.float64), dev(vel, torch.float64),
                           mass=dev(m, torch.float64))
    sim.step(nsteps)
    rho_s = sim.rho_s.clone()
    st = sim.state()
    return rho_s, st["psi"].clone(), st["pos"].clone(), st["vel"].clone(), sim.V.clone()


@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("g", [O.Grid(32, 32, 32), O.Grid(128, 64, 32, 2.0, 1.0, 0.5)],
                         ids=["32", "aniso"])
def test_deterministic_deposit_reproduces_bitwise(qj, cuda, dtype, g):
    psi = O.filtered_gaussian_psi(g, 61, sigma_k=8 * np.pi)
    pos, vel, m_s = O.star_ic(g, 20000, 1061, sigma_v=0.2)
    m = m_s * (0.5 + np.random.default_rng(6).random(20000))
    a = _run(qj, g, dtype, psi, pos, vel, m, 1)
    b = _run(qj, g, dtype, psi, pos, vel, m, 1)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
    # the float-atomic binned run agrees to rounding (c64: the FP32 sums of
    # about five corner values per cell, carried through three steps of
    # strong gravity, 1.7e-6 measured at 32^3)
    c = _run(qj, g, dtype, psi, pos, vel, m, 0)
    tol = {C64: 1e-5, C128: 1e-13}[dtype]
    for x, y in zip(a, c):
        assert rel(x, y) < tol
    # and the deposit of the final positions against the oracle
    ref = O.bin_particles(a[2].cpu().numpy(), m, g)
    assert rel(a[0].double().cpu(), torch.as_tensor(ref)) < {C64: 2e-7, C128: 1e-14}[dtype]


def test_deterministic_ensemble_matches_member_stepping_bitwise(qj, cuda):
    _lib = importlib.import_module("quantum-jax_amd._lib")
    specs = [(16, 1, 2e-3), (16, 2, 7e-4), (32, 3, 1e-3), (32, 4, 4e-4)]

    def members():
        out = []
        for n, seed, dt in specs:
            g = O.Grid(n, n, n)
            plan = qj.Plan(n, dtype=C64)
            plan.set_option(_lib.OPT_DETERMINISTIC_DEPOSIT, 1)
            pos, vel, m_s = O.star_ic(g, 500, seed + 1000, sigma_v=0.2)
            out.append(qj.KDKSimulation(plan, qj.Physics(rho_bar=1.1, four_pi_G=30.0), dt,
                                        dev(O.filtered_gaussian_psi(g, seed, sigma_k=6 * np.pi),
                                            C64),
                                        dev(pos, torch.float64), dev(vel, torch.float64),
                                        mass=m_s))
        return out

    ref, mem = members(), members()
    ens = qj.Ensemble(mem, mode="graph")
    for _ in range(4):
        ens.step(1)
        for s in ref:
            s.step(1)
        torch.cuda.synchronize()
    ens.finalize()
    for x, y in zip(mem, ref):
        y.finalize()
        assert torch.equal(x._psi, y._psi)
        assert torch.equal(x.pos, y.pos)
        assert torch.equal(x.vel, y.vel)
        # V and the acceleration grid (computed from the state when the fused
        # star gather kept only the stars' corner values)
        assert torch.equal(x.V, y.V)
        assert torch.equal(x.a_grid, y.a_grid)
    ens.close()
