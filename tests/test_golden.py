"""Committed regression fixtures (tests/golden/, written by make_golden.py).

The fixtures are FP64 oracle outputs on small seeded inputs (the reference
itself cannot run here).  The CPU test recomputes them with the oracle; the
GPU tests run the HIP path (complex128 / FP64) on the fixture inputs.
"""

import importlib.util
import os

import numpy as np
import pytest
import torch

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["plain_16", "stars_8x16x32", "cic_16", "hydro_16"]


def load(name):
    with np.load(os.path.join(HERE, f"{name}.npz")) as z:
        inp = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
        out = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
    return inp, out


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def test_oracle_reproduces_fixtures():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    for name, (_, exp) in mg.cases().items():
        _, fix = load(name)
        for k, v in exp.items():
            if name == "cic_16" and k in ("i", "ip1"):
                assert np.array_equal(fix[k], v), (name, k)
            else:
                assert rel(v, fix[k]) <= 1e-13, (name, k)


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda", dtype)


@pytest.mark.gpu
def test_gpu_plain_fixture(qj, cuda):
    inp, out = load("plain_16")
    plan = qj.Plan(16, dtype=torch.complex128)
    phys = qj.Physics(m_per_hbar=1.0, four_pi_G=1.0, rho_bar=1.0)
    sim = qj.KDKSimulation(plan, phys, float(inp["dt"]), dev(inp["psi"], torch.complex128))
    sim.step(int(inp["steps"]))
    psi = sim.psi
    assert rel(psi.cpu().numpy(), out["psi"]) < 1e-11
    V = qj.get_potential(psi.abs() ** 2, plan, phys)
    assert rel(V.cpu().numpy(), out["V"]) < 1e-11


@pytest.mark.gpu
def test_gpu_stars_fixture(qj, cuda):
    inp, out = load("stars_8x16x32")
    L = tuple(float(v) for v in inp["L"])
    plan = qj.Plan(8, 16, 32, L=L, dtype=torch.complex128)
    phys = qj.Physics(rho_bar=float(inp["rho_bar"]), four_pi_G=float(inp["four_pi_G"]))
    sim = qj.KDKSimulation(plan, phys, float(inp["dt"]), dev(inp["psi"], torch.complex128),
                           dev(inp["pos"], torch.float64), dev(inp["vel"], torch.float64),
                           mass=dev(inp["mass"], torch.float64))
    sim.step(int(inp["steps"]))
    st = sim.state()
    assert rel(st["psi"].cpu().numpy(), out["psi"]) < 1e-11
    assert rel(st["pos"].cpu().numpy(), out["pos"]) < 1e-12
    assert rel(st["vel"].cpu().numpy(), out["vel"]) < 1e-10
    assert st["a_max"] == pytest.approx(float(out["a_max"]), rel=1e-10)
    assert st["dt_s"] == pytest.approx(float(out["dt_s"]), rel=1e-10)


@pytest.mark.gpu
def test_gpu_cic_fixture_bit_exact(qj, cuda):
    inp, out = load("cic_16")
    plan = qj.Plan(16, dtype=torch.complex128)
    i, ip1, wi, wip1 = qj.get_cic_indices_and_weights(dev(inp["pos"], torch.float64), plan)
    assert np.array_equal(i.cpu().numpy(), out["i"])
    assert np.array_equal(ip1.cpu().numpy(), out["ip1"])
    assert np.array_equal(wi.cpu().numpy(), out["w_i"])
    assert np.array_equal(wip1.cpu().numpy(), out["w_ip1"])


@pytest.mark.gpu
def test_gpu_hydro_fixture(qj, cuda):
    inp, out = load("hydro_16")
    plan = qj.Plan(16, dtype=torch.complex128)
    phys = qj.Physics(cs=float(inp["cs"]))
    r = qj.solve_hydro(*(dev(inp[k], torch.float64) for k in ("rho", "vx", "vy", "vz")),
                       float(inp["dt"]), plan, phys)
    for k, t in zip(("rho", "vx", "vy", "vz"), r):
        assert rel(t.cpu().numpy(), out[k]) < 1e-12, k
