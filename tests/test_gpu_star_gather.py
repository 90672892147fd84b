"""The fused star gather (QJ_OPT_STAR_GATHER): the Poisson x pass hands each
star its eight corner accelerations instead of writing the acceleration
grid.  The stars then kick with the same corner values, so stars steps with
the option on agree with the FP64 oracle and with the option off to FP
rounding (the two Poisson x instantiations may round a transform's last bit
differently); the grid-valid state moves between the two forms in either
direction; a_grid of a gathered state is computed from the state.

quantum-stars.py:250-288 (get_acceleration: grid, then the CIC gather) and
:295-330 (the step whose kicks use it).
"""

import importlib

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128
TOLS = {C64: 2e-5, C128: 1e-11}


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda", dtype)


def _lib():
    return importlib.import_module("quantum-jax_amd._lib")


# 8^3 and 16^3: one thread per x pencil (the gather's own LDS buffer); 64 x
# 16 x 32: column tiles spanning several y rows; 256^3 c64: the wide Poisson
# x form with 8192-cell column tiles; 512 x 16 x 16: 512-point pencils
GRIDS = [O.Grid(8, 8, 8), O.Grid(16, 16, 16), O.Grid(64, 16, 32, 1.0, 0.5, 2.0),
         O.Grid(32, 32, 32), O.Grid(512, 16, 16, 2.0, 1.0, 1.0)]


def _sim(qj, g, dtype, n_p, gather, seed=3, det=0, wide=1):
    plan = qj.Plan(g.nx, g.ny, g.nz, L=(g.Lx, g.Ly, g.Lz), dtype=dtype)
    plan.set_option(_lib().OPT_STAR_GATHER, gather)
    plan.set_option(_lib().OPT_DETERMINISTIC_DEPOSIT, det)
    plan.set_option(_lib().OPT_WIDE_SWEEPS, wide)
    psi = O.filtered_gaussian_psi(g, seed, sigma_k=6 * np.pi)
    pos, vel, m_s = O.star_ic(g, n_p, seed + 1000, sigma_v=0.2)
    sim = qj.KDKSimulation(plan, qj.Physics(rho_bar=1.1, four_pi_G=20.0), 1e-3, dev(psi, dtype),
                           dev(pos, torch.float64), dev(vel, torch.float64), mass=m_s)
    return sim, (psi, pos, vel, m_s)


@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("g", GRIDS, ids=lambda g: f"{g.nx}x{g.ny}x{g.nz}")
def test_gathered_steps_match_oracle_and_grid_steps(qj, cuda, dtype, g):
    n_p = 3000
    res = {}
    for gather in (1, 0):
        sim, ic = _sim(qj, g, dtype, n_p, gather)
        sim.step(3)
        assert sim._st.grid_valid == (2 if gather else 1)
        st = sim.state()
        res[gather] = (st["psi"].clone(), st["pos"].clone(), st["vel"].clone(), st["a_max"],
                       st["dt_s"])
    psi, pos, vel, m_s = ic
    s = (psi, pos, vel, 0.0, 0.0, np.inf)
    for _ in range(3):
        s = O.compute_step_stars(*s, 1e-3, m_s, g, O.Physics(rho_bar=1.1, four_pi_G=20.0))
    for k, ref in ((0, s[0]), (1, s[1]), (2, s[2])):
        assert rel(res[1][k], ref) < TOLS[dtype], k
        assert rel(res[1][k], res[0][k]) < TOLS[dtype] / 10, k
    assert abs(res[1][3] - s[4]) < 1e-5 * s[4]
    assert abs(res[1][4] - s[5]) < 1e-5 * s[5]


@pytest.mark.parametrize("wide", [1, 0])
def test_gathered_steps_at_256(qj, cuda, wide):
    # 8192-cell column tiles, 2048 Poisson x workgroups; the deterministic
    # deposit, so both runs see the same densities up to the transforms'
    # rounding
    g = O.Grid(256, 256, 256)
    res = {}
    for gather in (1, 0):
        sim, _ = _sim(qj, g, C64, 1 << 16, gather, det=1, wide=wide)
        sim.step(2)
        st = sim.state()
        res[gather] = (st["psi"], st["pos"], st["vel"], sim.rho_s)
    for a, b in zip(res[1], res[0]):
        assert rel(a, b) < 1e-6


@pytest.mark.parametrize("dtype", [C64, C128])
def test_option_switches_between_steps(qj, cuda, dtype):
    # on -> off -> on: the state kept only corner values (grid_valid 2), then
    # a step reads them and writes the grid again (1), then gathers again
    g = O.Grid(32, 16, 64, 1.0, 0.5, 2.0)
    sims = []
    for sched in ((1, 0, 1), (0, 0, 0)):
        sim, _ = _sim(qj, g, dtype, 2000, sched[0])
        for gth in sched:
            sim.plan.set_option(_lib().OPT_STAR_GATHER, gth)
            sim.step(2)
            assert sim._st.grid_valid == (2 if gth else 1)
        sims.append(sim)
    a, b = (s.state() for s in sims)
    for k in ("psi", "pos", "vel"):
        assert rel(a[k], b[k]) < TOLS[dtype] / 10, k


@pytest.mark.parametrize("dtype", [C64, C128])
def test_a_grid_of_gathered_state(qj, cuda, dtype):
    g = O.Grid(32, 32, 32)
    out = []
    for gather in (1, 0):
        sim, _ = _sim(qj, g, dtype, 1500, gather)
        sim.step(2)
        out.append((sim.V.clone(), sim.a_grid.clone()))
    assert rel(out[0][0], out[1][0]) < TOLS[dtype] / 10
    assert rel(out[0][1], out[1][1]) < TOLS[dtype] / 10


def test_atomic_deposit_after_gathered_steps(qj, cuda):
    # the float-atomic star step reads the grid: a gathered state has its grid
    # computed again before such a step
    g = O.Grid(32, 32, 32)
    res = []
    for sched in ((1, 0), (0, 0)):
        sim, _ = _sim(qj, g, C128, 1500, sched[0])
        sim.step(2)
        sim.plan.set_option(_lib().OPT_BINNED_DEPOSIT, 0)
        sim.plan.set_option(_lib().OPT_STAR_GATHER, sched[1])
        sim.step(2)
        assert sim._st.grid_valid == 1
        res.append(sim.state())
    for k in ("psi", "pos", "vel"):
        assert rel(res[0][k], res[1][k]) < 1e-12, k


@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("g", [O.Grid(32, 32, 32), O.Grid(64, 16, 32, 1.0, 0.5, 2.0),
                               O.Grid(256, 32, 64)], ids=lambda g: f"{g.nx}x{g.ny}x{g.nz}")
def test_column_tile_rows_agree(qj, cuda, dtype, g):
    # QJ_OPT_GATHER_ROWS: one- and two-row column tiles (and the automatic
    # choice) against the grid path
    res = {}
    for gather, rows in ((0, 0), (1, 0), (1, 1), (1, 2)):
        sim, _ = _sim(qj, g, dtype, 4000, gather)
        sim.plan.set_option(_lib().OPT_GATHER_ROWS, rows)
        sim.step(3)
        st = sim.state()
        res[(gather, rows)] = (st["psi"].clone(), st["pos"].clone(), st["vel"].clone())
    base = res[(0, 0)]
    for key, r in res.items():
        for a, b in zip(r, base):
            assert rel(a, b) < TOLS[dtype] / 10, key


# QJ_OPT_FUSED_DEPOSIT: the source pass deposits its column tile's stars,
# writing rho_s from its LDS tile (1) or leaving it to qj_star_density (2,
# the default), against the separate deposit kernel (0).  64 x 16 x 32:
# tiles of two y rows; 8^3: one tile smaller than the workgroup's columns;
# 512 x 16 x 16: 512-point pencils
@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("g", GRIDS + [O.Grid(128, 64, 32, 1.0, 0.5, 0.25)],
                         ids=lambda g: f"{g.nx}x{g.ny}x{g.nz}")
@pytest.mark.parametrize("fuse_x", [1, 0])
def test_fused_deposit_matches_tile_deposit(qj, cuda, dtype, g, fuse_x):
    L = _lib()
    res = {}
    for fd in (2, 1, 0):
        sim, ic = _sim(qj, g, dtype, 3000, 1)
        sim.plan.set_option(L.OPT_FUSED_DEPOSIT, fd)
        sim.plan.set_option(L.OPT_FUSE_X_SOURCE, fuse_x)
        sim.step(3)
        assert sim._st.rho_valid == (0 if fd == 2 else 1)
        st = sim.state()
        res[fd] = (st["psi"].clone(), st["pos"].clone(), st["vel"].clone(), sim.rho_s.clone())
        assert sim._st.rho_valid == 1
    for fd in (2, 1):
        for k in range(4):
            assert rel(res[fd][k], res[0][k]) < TOLS[dtype] / 10, (fd, k)
    # rho_s is the CIC deposit of the positions after the step (the oracle's
    # bin_stars at those positions)
    m_s = ic[3]
    ref = O.bin_stars(res[1][1].cpu().numpy(), m_s, g)
    assert rel(res[1][3], ref) < (1e-6 if dtype == C64 else 1e-12)


def test_fused_deposit_star_masses(qj, cuda):
    # per-star masses (the mass array the adjoint differentiates)
    g = O.Grid(32, 32, 32)
    L = _lib()
    psi = O.filtered_gaussian_psi(g, 3, sigma_k=6 * np.pi)
    pos, vel, m_s = O.star_ic(g, 2500, 1003, sigma_v=0.2)
    m = m_s * (0.5 + np.random.default_rng(5).random(2500))
    out = []
    for fd in (1, 0, 2):
        plan = qj.Plan(32, dtype=C128)
        plan.set_option(L.OPT_FUSED_DEPOSIT, fd)
        sim = qj.KDKSimulation(plan, qj.Physics(rho_bar=1.1, four_pi_G=20.0), 1e-3,
                               dev(psi, C128), dev(pos, torch.float64), dev(vel, torch.float64),
                               mass=dev(m, torch.float64))
        sim.step(3)
        st = sim.state()
        out.append((st["psi"].clone(), st["pos"].clone(), sim.rho_s.clone()))
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert rel(a, b) < 1e-12
    ref = O.bin_particles(out[0][1].cpu().numpy(), m, g)
    assert rel(out[0][2], ref) < 1e-12


@pytest.mark.parametrize("dtype", [C64, C128])
@pytest.mark.parametrize("rows", [1, 2])
def test_fused_deposit_on_tile_edges(qj, cuda, dtype, rows):
    # stars whose CIC clouds straddle column-tile edges in y and z, wrap
    # around the periodic faces, or sit on cell centres (a zero-weight
    # corner), each in cells of its own: the fused and the separate deposit
    # then add one value per cell and agree bitwise, and both equal the
    # oracle's deposit (and so does the density qj_star_density computes for
    # the default form).  No gravity and no velocity, so the stars stay put.
    L = _lib()
    n = 64
    g = O.Grid(n, n, n)
    dx = g.Lx / n
    c = lambda i: (i + 0.5) * dx  # cell centre
    pos = np.array([
        [c(3), 2 * dx, c(9)],            # y face of cells 1 | 2 (a two-row tile edge)
        [c(20), c(40), 32 * dx],         # z face of cells 31 | 32
        [c(30), 16 * dx, 48 * dx],       # both
        [c(40), 0.1 * dx, c(20)],        # y wraps (cells 63 | 0)
        [c(50), c(10), n * dx - 0.25 * dx],  # z wraps
        [0.2 * dx, c(50), c(50)],        # x wraps
        [c(12), c(25), c(5)],            # on a cell centre
        [10 * dx, 34 * dx, 16 * dx],     # faces on all three axes
        [c(60), 62.7 * dx, 63.9 * dx],   # near the far corner
        [c(5), 53.5 * dx, 30.5 * dx],
    ])
    m = np.linspace(0.5, 1.5, len(pos)) * 1e-4
    psi = O.filtered_gaussian_psi(g, 8, sigma_k=6 * np.pi)
    out = []
    for fd in (1, 0, 2):
        plan = qj.Plan(n, dtype=dtype)
        plan.set_option(L.OPT_FUSED_DEPOSIT, fd)
        plan.set_option(L.OPT_GATHER_ROWS, rows)
        sim = qj.KDKSimulation(plan, qj.Physics(rho_bar=1.0, four_pi_G=0.0), 1e-3, dev(psi, dtype),
                               dev(pos, torch.float64), dev(np.zeros_like(pos), torch.float64),
                               mass=dev(m, torch.float64))
        sim.step(2)
        assert torch.equal(sim.pos.cpu(), torch.as_tensor(pos))
        out.append(sim.rho_s.clone())
    assert torch.equal(out[0], out[1])
    ref = O.bin_particles(pos, m, g)
    # (2: qj_star_density, the fixed-point deposit, rounds each value twice)
    for o in out:
        assert rel(o, ref) < (1e-7 if dtype == C64 else 1e-15)


@pytest.mark.parametrize("dtype", [C64, C128])
def test_reading_rho_s_between_steps_changes_nothing(qj, cuda, dtype):
    # qj_star_density (the density of a state whose fused-deposit steps did
    # not write it) registers the stars again in the star scratch; the
    # gather's corner values there survive it, and the next steps register
    # their own, so a run that reads rho_s between its steps is the run that
    # does not (to the rounding by which any two runs differ: the deposit's
    # float atomics add in the order the counting atomics gave)
    g = O.Grid(64, 32, 32, 1.0, 0.5, 0.5)
    runs = []
    for read in (True, False):
        sim, _ = _sim(qj, g, dtype, 5000, 1)
        for _ in range(3):
            sim.step(2)
            assert sim._st.grid_valid == 2
            if read:
                assert sim._st.rho_valid == 0
                r = sim.rho_s
                assert sim._st.rho_valid == 1
                assert torch.isfinite(r).all()
        st = sim.state()
        runs.append((st["psi"].clone(), st["pos"].clone(), st["vel"].clone()))
    for a, b in zip(*runs):
        assert rel(a, b) < TOLS[dtype] / 10
