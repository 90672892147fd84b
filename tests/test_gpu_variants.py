"""Kick/drift variants against the FP64 oracle (needs a GPU).

quantum-tidal.py: a static external potential (softened point mass) added
to V in the psi kicks; quantum-cosmo.py: the comoving step, whose kicks use
dt pot_fac and drift dt kin_fac with dt from the quadrature of 1/adot, run
through the per-step schedule entry point (merged kicks);
quantum-fluid-accretion.py: gas steps with the hydro CFL cap as dt.
"""

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128
TOL = {C64: 2e-5, C128: 1e-11}


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda", dtype)


@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("n", [16, 32])
def test_external_potential_steps_match_oracle(qj, cuda, dtype, n):
    g = O.Grid(n, n, n)
    ph = O.Physics(four_pi_G=10.0, m_per_hbar=1.0, rho_bar=1.0)
    psi = O.filtered_gaussian_psi(g, 9, sigma_k=6 * np.pi)
    M, G, dt, nsteps = 3.0, 0.8, 1e-3, 5
    Vx = O.external_point_mass(g, M, G)
    p = psi
    for _ in range(nsteps):
        p = O.compute_step_tidal(p, dt, Vx, g, ph)
    plan = qj.Plan(n, dtype=dtype)
    v_ext = qj.cosmo.point_mass_potential(plan, M, G)
    assert rel(v_ext, Vx) < (1e-7 if dtype == C64 else 1e-15)
    sim = qj.KDKSimulation(plan, qj.Physics(four_pi_G=10.0, rho_bar=1.0), dt,
                           dev(psi, dtype), v_ext=v_ext)
    sim.step(nsteps)
    assert rel(sim.psi, p) < TOL[dtype]


@pytest.mark.parametrize("dtype", [C64, C128])
def test_comoving_schedule_matches_oracle(qj, cuda, dtype):
    g = O.Grid(32, 32, 32)
    ph = O.Physics(four_pi_G=10.0, m_per_hbar=1.0, rho_bar=1.0)
    psi = O.filtered_gaussian_psi(g, 10, sigma_k=6 * np.pi)
    a0, da, nsteps, H0, om, ol = 0.1, 2e-3, 6, 0.5, 0.3, 0.7
    p, a = psi, a0
    for _ in range(nsteps):
        p, a = O.compute_step_cosmo(p, a, da, g, ph, H0, om, ol)
    kick, drift, a_end = qj.cosmo.comoving_schedule(a0, da, nsteps, H0, om, ol)
    assert a_end == pytest.approx(a)
    plan = qj.Plan(32, dtype=dtype)
    sim = qj.KDKSimulation(plan, qj.Physics(four_pi_G=10.0, rho_bar=1.0), 0.0, dev(psi, dtype))
    sim.step_schedule(kick[:2], drift[:2])
    sim.step_schedule(kick[2:], drift[2:])  # the owed kick carries across calls
    assert rel(sim.psi, p) < TOL[dtype]


def test_variants_reject_stars(qj, cuda):
    g = O.Grid(16, 16, 16)
    plan = qj.Plan(16)
    psi = dev(O.filtered_gaussian_psi(g, 1), C64)
    pos, vel, m = O.star_ic(g, 10, 2)
    sim = qj.KDKSimulation(plan, qj.Physics(), 1e-3, psi, dev(pos, torch.float64),
                           dev(vel, torch.float64), mass=m,
                           v_ext=torch.zeros(g.shape, dtype=torch.float32, device="cuda"))
    with pytest.raises(qj.QjError):
        sim.step(1)
    sim2 = qj.KDKSimulation(qj.Plan(16), qj.Physics(), 1e-3, psi, dev(pos, torch.float64),
                            dev(vel, torch.float64), mass=m)
    with pytest.raises(qj.QjError):
        sim2.step_schedule([1e-3], [1e-3])


@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("flux", ["rusanov", "hll"])
def test_cfl_limited_gas_steps_match_oracle(qj, cuda, dtype, flux):
    """Gas steps whose dt is the CFL cap of the state they start from
    (quantum-fluid-accretion.py:696-701), through the schedule entry point;
    the cap binds (cs + |v| ~ 1.5, dx = 1/32) and changes from step to step."""
    g = O.Grid(32, 32, 32)
    psi = O.filtered_gaussian_psi(g, 27, sigma_k=8 * np.pi)
    rho, _, _, _ = O.gas_ic(g, 27, sigma_k=8 * np.pi)
    rng = np.random.default_rng(28)
    vx, vy, vz = (0.3 * rng.standard_normal(g.shape) for _ in range(3))
    cs, cfl, dt_max = 1.0, 0.4, 1e-2
    rb = float(np.mean(np.abs(psi) ** 2 + rho))
    r = torch.float32 if dtype == C64 else torch.float64
    gas = [dev(a, r) for a in (rho, vx, vy, vz)]
    sim = qj.KDKSimulation(qj.Plan(32, dtype=dtype), qj.Physics(rho_bar=rb, cs=cs, flux=flux),
                           dt_max, dev(psi, dtype), gas=gas)
    s = (psi, *[a.double().cpu().numpy() for a in gas], 0.0)
    ph = O.Physics(rho_bar=rb, cs=cs, flux=flux)
    dts = []
    for _ in range(4):
        dt = O.cfl_dt(s[2], s[3], s[4], cs, g.dx, cfl, dt_max)
        dts.append(dt)
        s = O.compute_step_gas(*s, dt, g, ph)
    sim.step_cfl(4, cfl=cfl, dt_max=dt_max)
    assert dts[0] < dt_max and len(set(dts)) == 4
    assert np.allclose(sim.dts, dts, rtol=(1e-6 if dtype == C64 else 1e-13), atol=0)
    assert abs(sim.t - sum(sim.dts)) < 1e-15
    st = sim.state()
    assert rel(st["psi"], s[0]) < TOL[dtype]
    for k, name in enumerate(("rho", "vx", "vy", "vz")):
        assert rel(st[name], s[1 + k]) < (5e-5 if dtype == C64 else 1e-11), name
    # after finalize (no owed kick) the cap is that of the finalized state
    want = O.cfl_dt(*[st[k].double().cpu().numpy() for k in ("vx", "vy", "vz")], cs, g.dx, cfl,
                    dt_max)
    assert abs(sim.cfl_dt(cfl, dt_max) / want - 1) < (1e-6 if dtype == C64 else 1e-13)


@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("n,wide", [(32, 1), (256, 1), (256, 0)])
def test_cfl_signal_fold_equals_signal_pass(qj, cuda, dtype, n, wide):
    """Within one CFL call every step after the first takes its signal speed
    from the previous step's Poisson x pass; a call of one step computes it
    with the separate pass over v and the grid.  Both routes give the same
    bits: one call of 5 steps equals 5 calls of one step (dts and states)."""
    g = O.Grid(n, n, n)
    psi = O.filtered_gaussian_psi(g, 31, sigma_k=8 * np.pi)
    rho, _, _, _ = O.gas_ic(g, 31, sigma_k=8 * np.pi)
    rng = np.random.default_rng(32)
    vel = [0.3 * rng.standard_normal(g.shape) for _ in range(3)]
    rb = float(np.mean(np.abs(psi) ** 2 + rho))
    r = torch.float32 if dtype == C64 else torch.float64
    runs = []
    for calls in (1, 5):
        plan = qj.Plan(n, dtype=dtype)
        plan.set_option(qj._lib.OPT_WIDE_SWEEPS, wide)
        sim = qj.KDKSimulation(plan, qj.Physics(rho_bar=rb, cs=1.0, flux="hll"), 1e-2,
                               dev(psi, dtype), gas=[dev(a, r) for a in (rho, *vel)])
        for _ in range(calls):
            sim.step_cfl(5 // calls, cfl=0.4)
        st = sim.state()
        runs.append((list(sim.dts), [st[k].clone() for k in ("psi", "rho", "vx", "vy", "vz")]))
    assert len(set(runs[0][0])) == 5   # the cap binds and changes each step
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1], runs[1][1]):
        assert torch.equal(a, b)


def test_cfl_rejects_states_without_gas(qj, cuda):
    sim = qj.KDKSimulation(qj.Plan(16), qj.Physics(), 1e-3,
                           dev(O.filtered_gaussian_psi(O.Grid(16, 16, 16), 29), C64))
    with pytest.raises(ValueError):
        sim.cfl_dt()
