"""The ensemble runner against member-by-member stepping (needs a GPU).

An ensemble step records each member's launches, batches members of the
same shape into one launch per stage and replays the step from a hipGraph.
The kernels and arguments are the ones qj_kdk_steps issues, so:
  * plain and gas members must match stepping each member alone bit for bit;
  * star members match to rounding only (the star deposit adds with float
    atomics, whose order varies from run to run), checked at 1e-6 (c64) and
    1e-12 (c128) relative L2 after several steps.
The unfinalized buffers (psi with its owed kick, grid, star and gas state)
are compared after every step, across graph replays with eager work and
host synchronisation between them.
"""

import importlib

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128
TOL = {C64: 1e-6, C128: 1e-12}


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda", dtype)


def rel(a, b):
    return ((a.double() - b.double()).norm() / b.double().norm().clamp_min(1e-300)).item() \
        if not a.is_complex() else ((a - b).abs().norm() / b.abs().norm()).item()


def make_member(qj, kind, n, dtype, seed, dt, n_p=200):
    g = O.Grid(n, n, n)
    plan = qj.Plan(n, dtype=dtype)
    psi = dev(O.filtered_gaussian_psi(g, seed, sigma_k=6 * np.pi), dtype)
    if kind == "stars":
        pos, vel, m_s = O.star_ic(g, n_p, seed + 1000, sigma_v=0.2)
        phys = qj.Physics(rho_bar=1.1, four_pi_G=30.0)
        return qj.KDKSimulation(plan, phys, dt, psi, dev(pos, torch.float64),
                                dev(vel, torch.float64), mass=m_s)
    if kind == "gas":
        rho, vx, vy, vz = O.gas_ic(g, seed)
        rt = torch.float32 if dtype == C64 else torch.float64
        rb = float(np.mean(np.abs(O.filtered_gaussian_psi(g, seed, sigma_k=6 * np.pi)) ** 2 + rho))
        phys = qj.Physics(rho_bar=rb, four_pi_G=20.0, cs=0.3)
        vx = vx + 0.05 * np.sin(2 * np.pi * (np.arange(n) + 0.5) / n)[:, None, None]
        return qj.KDKSimulation(plan, phys, dt, psi, gas=[dev(a, rt) for a in (rho, vx, vy, vz)])
    return qj.KDKSimulation(plan, qj.Physics(rho_bar=1.0, four_pi_G=10.0), dt, psi)


SPECS = [  # kind, n, dtype, seed, dt
    ("stars", 16, C64, 1, 2e-3), ("stars", 16, C64, 2, 7e-4), ("stars", 16, C64, 3, 1.3e-3),
    ("stars", 32, C64, 4, 1e-3), ("stars", 32, C64, 5, 4e-4),
    ("plain", 16, C128, 6, 1e-3), ("plain", 32, C64, 7, 9e-4), ("plain", 32, C64, 8, 3e-4),
    ("gas", 16, C64, 9, 1e-3), ("gas", 16, C64, 10, 6e-4),
    ("stars", 16, C128, 11, 1e-3),
]


def build(qj, specs=SPECS):
    return [make_member(qj, *s) for s in specs]


def compare(specs, a, b, step):
    for spec, x, y in zip(specs, a, b):
        kind, dtype = spec[0], spec[2]
        tag = f"{spec} after step {step}"
        if kind == "stars":
            assert rel(x._psi, y._psi) < TOL[dtype], tag
            assert rel(x.pos, y.pos) < TOL[dtype], tag
            assert rel(x.vel, y.vel) < TOL[dtype] * 10, tag
        else:
            assert torch.equal(x._psi, y._psi), tag
            # plain steps write V only; gas steps also the acceleration planes
            assert torch.equal(x.V, y.V), tag
            if kind == "gas":
                assert torch.equal(x.a_grid, y.a_grid), tag
            if kind == "gas":
                assert x._st.gas_swapped == y._st.gas_swapped
                for u, v in zip(x.gas + x._gas_tmp, y.gas + y._gas_tmp):
                    assert torch.equal(u, v), tag
        assert x._st.pending == y._st.pending and x._st.grid_valid == y._st.grid_valid


@pytest.mark.parametrize("mode", ["eager", "batched", "graph", "graph_fork"])
def test_ensemble_matches_member_stepping(qj, cuda, mode):
    ref = build(qj)
    mem = build(qj)
    ens = qj.Ensemble(mem, mode=mode)
    for k in range(1, 6):
        for s in ref:
            s.step(1)
        ens.step(1)
        torch.cuda.synchronize()
        _ = (ref[0]._psi.abs() ** 2).sum().item()  # eager work between replays
        compare(SPECS, mem, ref, k)
    ens.step(3)
    for s in ref:
        s.step(3)
    ens.finalize()
    for s in ref:
        s.finalize()
    compare(SPECS, mem, ref, 8)
    for x, y in zip(mem, ref):
        assert x.a_max() == pytest.approx(y.a_max(), rel=1e-5)
    ens.close()


def test_ensemble_groups_members_by_shape(qj, cuda):
    mem = build(qj)
    ens = qj.Ensemble(mem, mode="batched")
    ens.step(2)
    groups = sorted(tuple(m) for m, _ in ens.groups())
    want = {}
    for i, s in enumerate(SPECS):
        want.setdefault((s[0], s[1], s[2]), []).append(i)
    assert groups == sorted(tuple(v) for v in want.values())
    # one launch per stage and group, independent of the member count
    launches = {tuple(m): nl for m, nl in ens.groups()}
    assert launches[(0, 1, 2)] == launches[(3, 4)]
    ens.close()


def test_ensemble_finalize_then_continue(qj, cuda):
    # finalize zeroes the owed kicks; the next step records the opening step again
    specs = SPECS[:5] + SPECS[8:10]
    ref = build(qj, specs)
    mem = build(qj, specs)
    ens = qj.Ensemble(mem, mode="graph")
    for rnd in range(3):
        ens.step(2)
        ens.finalize()
        for s in ref:
            s.step(2)
            s.finalize()
        compare(specs, mem, ref, 2 * (rnd + 1))
    ens.close()


def test_autotune_leaves_state_and_results_unchanged(qj, cuda):
    specs = SPECS[:8]
    ref = build(qj, specs)
    mem = build(qj, specs)
    ens = qj.Ensemble(mem, mode="graph")
    ens.step(2)
    for s in ref:
        s.step(2)
    res = ens.autotune(reps=2)
    assert res["mode"] in ("batched", "graph", "graph_fork")
    assert set(res["fuse_x_source"].values()) <= {0, 1, 2}
    compare(specs, mem, ref, 2)
    # the members continue under the tuned schedule options, the references
    # under the same ones (the x schedules agree bitwise in complex64 only)
    _lib = importlib.import_module("quantum-jax_amd._lib")
    for x, y in zip(mem, ref):
        for o in range(_lib.OPT_COUNT):
            y.plan.set_option(o, x.plan.get_option(o))
    ens.step(3)
    for s in ref:
        s.step(3)
    compare(specs, mem, ref, 5)
    # a second ensemble of the same signature takes the tuned schedule from the cache
    ens2 = qj.Ensemble(build(qj, specs))
    assert ens2.autotune() is res
    ens.close()
    ens2.close()


def test_ensemble_rejects_shared_plans(qj, cuda):
    plan = qj.Plan(16)
    psi = torch.ones((16, 16, 16), dtype=C64, device="cuda")
    a = qj.KDKSimulation(plan, qj.Physics(), 1e-3, psi)
    b = qj.KDKSimulation(plan, qj.Physics(), 1e-3, psi)
    with pytest.raises(ValueError):
        qj.Ensemble([a, b])


@pytest.mark.parametrize("mode", ["graph", "eager"])
def test_ensemble_adjoint_matches_member_adjoints(qj, cuda, mode):
    # gradients of sum_b L_b through the batched ensemble forward + adjoint
    # equal each member's own kdk_run gradients (complex128; the float-atomic
    # paths differ from run to run by rounding only)
    import importlib
    ad = importlib.import_module("quantum-jax_amd.adjoint")
    specs = [(16, 1, 2e-3), (16, 2, 1.1e-3), (32, 3, 1.5e-3)]
    nsteps, seg = 5, 2
    rng = np.random.default_rng(0)
    inputs, targets = [], []
    for n, seed, dt in specs:
        g = O.Grid(n, n, n)
        psi = dev(O.filtered_gaussian_psi(g, seed, sigma_k=6 * np.pi), C128)
        pos, vel, m_s = O.star_ic(g, 150, seed + 1000, sigma_v=0.2)
        m = m_s * (0.5 + rng.random(150))
        inputs.append((psi, dev(pos, torch.float64), dev(vel, torch.float64),
                       dev(m, torch.float64)))
        targets.append(dev(np.abs(O.filtered_gaussian_psi(g, seed + 50, sigma_k=6 * np.pi)) ** 2,
                           torch.float64))
    phys = qj.Physics(rho_bar=1.1, four_pi_G=20.0)

    def leaf(t):
        return t.clone().requires_grad_(True)

    # member by member
    ref = []
    for (n, _, dt), (psi, pos, vel, m), tgt in zip(specs, inputs, targets):
        plan = qj.Plan(n, dtype=C128)
        p, v, mm = leaf(psi), leaf(vel), leaf(m)
        out = ad.kdk_run(p, pos, v, mm, plan, phys, dt, nsteps, seg)
        ad.density_loss(out[0], out[1], mm, tgt, plan).backward()
        ref.append((p.grad, v.grad, mm.grad))
    # one ensemble
    plans = [qj.Plan(n, dtype=C128) for n, _, _ in specs]
    ps = [leaf(i[0]) for i in inputs]
    vs = [leaf(i[2]) for i in inputs]
    ms = [leaf(i[3]) for i in inputs]
    xs = [i[1] for i in inputs]
    po, xo, vo = ad.ensemble_kdk_run(ps, xs, vs, ms, plans, [phys] * 3, [s[2] for s in specs],
                                     nsteps, seg, mode=mode)
    L = sum(ad.density_loss(po[b], xo[b], ms[b], targets[b], plans[b]) for b in range(3))
    L.backward()
    rel = lambda a, b: ((a - b).abs().norm() / b.abs().norm()).item()
    for b in range(3):
        assert rel(ps[b].grad, ref[b][0]) < 1e-10, b
        assert rel(vs[b].grad, ref[b][1]) < 1e-10, b
        assert rel(ms[b].grad, ref[b][2]) < 1e-10, b
