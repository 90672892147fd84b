"""Edge cases of the HIP path against the FP64 oracle (needs a GPU).

* the longest supported pencils (512 points) along each axis, and 256-point
  pencils along the strided axes, for the drift, the potential and the
  acceleration grid;
* a stars step on a flat, anisotropic box (tiles shrunk along z);
* empty star sets (deposit, interpolation, acceleration, stepper);
* non-finite star positions (skipped by the deposit, NaN accelerations, the
  rest of the run unaffected, dt_s over the finite velocities).
Rejected plan shapes are checked without a GPU in test_boundary.py.
"""

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128
TOL1 = {C64: 4e-6, C128: 1e-12}


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda", dtype)


def rdt(dtype):
    return torch.float32 if dtype == C64 else torch.float64


def plan_for(qj, g, dtype):
    return qj.Plan(g.nx, g.ny, g.nz, L=(g.Lx, g.Ly, g.Lz), dtype=dtype)


LONG = [O.Grid(512, 8, 16), O.Grid(8, 512, 16, 1.0, 4.0, 1.0), O.Grid(16, 8, 512, 1.0, 1.0, 2.0),
        O.Grid(256, 16, 8, 2.0, 1.0, 1.0), O.Grid(8, 256, 16)]


@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("g", LONG, ids=lambda g: f"{g.nx}x{g.ny}x{g.nz}")
def test_long_pencils_match_oracle(qj, cuda, dtype, g):
    psi = O.filtered_gaussian_psi(g, 21, sigma_k=8 * np.pi)
    plan = plan_for(qj, g, dtype)
    out = qj.drift(dev(psi, dtype), 1e-3, plan, qj.Physics(m_per_hbar=1.3))
    assert rel(out, O.drift(psi, 1e-3, g, O.Physics(m_per_hbar=1.3))) < TOL1[dtype]
    rho = np.abs(psi) ** 2
    ph = O.Physics(four_pi_G=2.0, rho_bar=1.3)
    V = qj.get_potential(dev(rho, rdt(dtype)), plan, qj.Physics(four_pi_G=2.0, rho_bar=1.3))
    assert rel(V, O.get_potential(rho, g, ph)) < 4 * TOL1[dtype]
    diag = plan.new_diag()
    _, ag = qj.acceleration_grid(dev(rho, rdt(dtype)), plan,
                                 qj.Physics(four_pi_G=2.0, rho_bar=1.3), diag=diag)
    ref = O.acceleration_grid(rho, g, ph)
    assert rel(ag.permute(3, 0, 1, 2), ref) < 4 * TOL1[dtype]
    amax = np.abs(ref).max()
    assert abs(diag[0].item() - amax) <= (1e-5 if dtype == C64 else 1e-12) * amax


@pytest.mark.parametrize("dtype", [C64, C128])
def test_stars_steps_on_a_flat_box(qj, cuda, dtype):
    # 8 cells along x: one tile row; 512 along z: 512-point contiguous pencils
    g = O.Grid(8, 32, 512, 0.25, 1.0, 16.0)
    psi = O.filtered_gaussian_psi(g, 22, sigma_k=2 * np.pi)
    pos, vel, m_s = O.star_ic(g, 3000, 1022, sigma_v=0.2)
    ph = O.Physics(rho_bar=1.1)
    dt, nsteps = 1e-3, 3
    plan = plan_for(qj, g, dtype)
    sim = qj.KDKSimulation(plan, qj.Physics(rho_bar=1.1), dt, dev(psi, dtype),
                           dev(pos, torch.float64), dev(vel, torch.float64), mass=m_s)
    sim.step(nsteps)
    st = sim.state()
    s = (psi, pos, vel, 0.0, 0.0, np.inf)
    for _ in range(nsteps):
        s = O.compute_step_stars(*s, dt, m_s, g, ph)
    assert rel(st["psi"], s[0]) < (2e-5 if dtype == C64 else 1e-11)
    L = np.array([g.Lx, g.Ly, g.Lz])
    dpos = (st["pos"].cpu().numpy() - s[1] + 0.5 * L) % L - 0.5 * L
    # position errors scale with the box (here up to 16 along z)
    assert np.max(np.abs(dpos)) < (1e-9 if dtype == C64 else 1e-13) * L.max()


@pytest.mark.parametrize("dtype", [C64, C128])
def test_empty_star_sets(qj, cuda, dtype):
    g = O.Grid(16, 16, 16)
    plan = plan_for(qj, g, dtype)
    none = torch.empty((0, 3), dtype=torch.float64, device="cuda")
    rho = qj.bin_particles(none, torch.empty(0, dtype=torch.float64, device="cuda"), plan)
    assert torch.count_nonzero(rho).item() == 0
    assert torch.count_nonzero(qj.bin_stars(none, 1.0, plan)).item() == 0
    ag = torch.ones(g.shape + (3,), dtype=rdt(dtype), device="cuda")
    assert qj.interp(none, ag, plan).shape == (0, 3)
    r = np.abs(O.filtered_gaussian_psi(g, 23, sigma_k=4 * np.pi)) ** 2
    acc, a_max = qj.get_acceleration(none, dev(r, rdt(dtype)), plan, qj.Physics())
    _, ramax = O.get_acceleration(np.zeros((0, 3)), r, g, O.Physics())
    assert acc.shape == (0, 3)
    assert abs(a_max.item() - ramax) <= 1e-5 * ramax
    # a stepper with an empty star set is the particle-free step
    psi = dev(O.filtered_gaussian_psi(g, 24, sigma_k=4 * np.pi), dtype)
    a = qj.KDKSimulation(plan, qj.Physics(rho_bar=1.0), 1e-3, psi, none, none.clone(), mass=0.1)
    b = qj.KDKSimulation(plan_for(qj, g, dtype), qj.Physics(rho_bar=1.0), 1e-3, psi)
    a.step(3)
    b.step(3)
    assert torch.equal(a.psi, b.psi)


@pytest.mark.parametrize("dtype", [C64, C128])
def test_nonfinite_star_positions(qj, cuda, dtype):
    g = O.Grid(16, 16, 32)
    rng = np.random.default_rng(25)
    pos = rng.random((400, 3)) * np.array([g.Lx, g.Ly, g.Lz])
    bad = [5, 77, 300]
    posb = pos.copy()
    posb[5] = [np.nan, 0.5, 0.5]
    posb[77] = [0.5, np.inf, 0.5]
    posb[300] = [0.5, 0.5, -np.inf]
    keep = np.setdiff1d(np.arange(400), bad)
    plan = plan_for(qj, g, dtype)
    m = rng.random(400)
    # the deposit skips non-finite positions
    rho = qj.bin_particles(dev(posb, torch.float64), dev(m, torch.float64), plan)
    ref = O.bin_particles(pos[keep], m[keep], g)
    assert rel(rho, ref) < (2e-6 if dtype == C64 else 1e-12)
    # the gather gives NaN rows for them and exact rows for the others
    f = rng.standard_normal(g.shape + (3,))
    acc = qj.interp(dev(posb, torch.float64), dev(f, rdt(dtype)), plan).cpu().numpy()
    assert np.isnan(acc[bad]).all()
    fr = dev(f, rdt(dtype)).double().cpu().numpy()
    assert rel(acc[keep], O.interp_grid(fr.transpose(3, 0, 1, 2), pos[keep], g)) < 1e-14
    # a run with those stars equals the run without them, for psi and the others
    psi = O.filtered_gaussian_psi(g, 26, sigma_k=4 * np.pi)
    vel = rng.standard_normal((400, 3)) * 0.1
    runs = []
    for p_, v_, m_ in ((posb, vel, m), (pos[keep], vel[keep], m[keep])):
        sim = qj.KDKSimulation(plan_for(qj, g, dtype), qj.Physics(rho_bar=1.0), 1e-3,
                               dev(psi, dtype), dev(p_, torch.float64), dev(v_, torch.float64),
                               mass=dev(m_, torch.float64))
        sim.step(3)
        runs.append(sim.state())
    tol = 1e-6 if dtype == C64 else 1e-12
    assert rel(runs[0]["psi"], runs[1]["psi"]) < tol
    assert rel(runs[0]["pos"][keep], runs[1]["pos"]) < (1e-10 if dtype == C64 else 1e-13)
    assert rel(runs[0]["vel"][keep], runs[1]["vel"]) < tol
    assert torch.isnan(runs[0]["vel"][bad]).all()
    # dt_s is the minimum over the finite velocities
    assert abs(runs[0]["dt_s"] / runs[1]["dt_s"] - 1) < tol


def test_largest_grid_512_cubed_is_time_reversible(qj, cuda):
    """512^3 complex64 with 2^16 stars (the largest pencils on every axis):
    steps with dt then with -dt return psi and the stars to their start
    (the merged-kick KDK is time-reversible up to rounding), and the mass
    sum(|psi|^2) vol is conserved."""
    n = 512
    g = torch.Generator(device="cuda").manual_seed(30)
    psi = torch.complex(torch.randn((n, n, n), generator=g, device="cuda"),
                        torch.randn((n, n, n), generator=g, device="cuda"))
    k = 2 * np.pi * torch.fft.fftfreq(n, 1.0 / n, device="cuda")
    filt = torch.exp(-(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)
                     / (2 * (32 * np.pi) ** 2))
    psi = torch.fft.ifftn(torch.fft.fftn(psi) * filt)
    psi = (psi / psi.abs().pow(2).mean().sqrt()).to(C64)
    del filt
    rng = np.random.default_rng(31)
    pos = dev(rng.random((1 << 16, 3)), torch.float64)
    vel = dev(0.05 * rng.standard_normal((1 << 16, 3)), torch.float64)
    plan = qj.Plan(n, dtype=C64)
    sim = qj.KDKSimulation(plan, qj.Physics(rho_bar=1.1), 2e-4, psi, pos, vel, mass=0.1 / (1 << 16))
    m0 = sim.psi.abs().pow(2).double().sum().item()
    sim.step(3)
    sim.finalize()
    moved = rel(sim.psi, psi)
    sim.set_dt(-2e-4)
    sim.step(3)
    st = sim.state()
    assert moved > 1e-3                          # the forward steps did change psi
    assert rel(st["psi"], psi) < 2e-5
    dpos = (st["pos"] - pos + 0.5) % 1.0 - 0.5
    assert dpos.abs().max().item() < 1e-9
    assert rel(st["vel"], vel) < 1e-6
    assert abs(st["psi"].abs().pow(2).double().sum().item() / m0 - 1) < 1e-5


def test_gas_256_cubed_hll_cfl_conserves_mass(qj, cuda):
    """The gas configuration at its full size (256^3 complex64 psi + FP32 gas,
    HLL flux, CFL 0.4): sum(|psi|^2 + rho_gas) vol is conserved by the KDK
    step (the flux form conserves gas mass exactly, the drift and kicks
    conserve |psi|^2), and the CFL steps stay within dt_max."""
    n = 256
    g = torch.Generator(device="cuda").manual_seed(32)
    psi = torch.complex(torch.randn((n, n, n), generator=g, device="cuda"),
                        torch.randn((n, n, n), generator=g, device="cuda"))
    psi = (psi / psi.abs().pow(2).mean().sqrt()).to(C64)
    rho = 0.2 * (1 + 0.1 * torch.randn((n, n, n), generator=g, device="cuda")).clamp_min(0.05)
    vel = [0.05 * torch.randn((n, n, n), generator=g, device="cuda") for _ in range(3)]
    rb = float(psi.abs().pow(2).mean().item() + rho.mean().item())
    sim = qj.KDKSimulation(qj.Plan(n, dtype=C64), qj.Physics(rho_bar=rb, cs=0.1, flux="hll"),
                           5e-4, psi, gas=[rho.float(), *[v.float() for v in vel]])
    m0 = psi.abs().pow(2).double().sum().item() + rho.double().sum().item()
    sim.step_cfl(4, cfl=0.4)
    st = sim.state()
    m1 = st["psi"].abs().pow(2).double().sum().item() + st["rho"].double().sum().item()
    assert abs(m1 / m0 - 1) < 1e-5          # FP32 rounding, about 2e-7 per step
    assert all(0 < d <= 5e-4 for d in sim.dts) and len(sim.dts) == 4
    assert torch.isfinite(st["vx"]).all()
