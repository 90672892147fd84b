"""Snapshot round trip of the stepper state (needs a GPU)."""

import importlib

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda", dtype)


@pytest.mark.parametrize("kind", ["plain", "stars", "gas"])
def test_checkpoint_round_trip_continues_the_run(qj, cuda, kind, tmp_path):
    ck = importlib.import_module("quantum-jax_amd.checkpoint")
    g = O.Grid(16, 16, 16)
    dtype = torch.complex128
    psi = dev(O.filtered_gaussian_psi(g, 3, sigma_k=6 * np.pi), dtype)
    phys = qj.Physics(rho_bar=1.1, four_pi_G=20.0)
    kw = {}
    if kind == "stars":
        pos, vel, m_s = O.star_ic(g, 100, 4, sigma_v=0.2)
        kw = dict(pos=dev(pos, torch.float64), vel=dev(vel, torch.float64), mass=m_s)
    if kind == "gas":
        kw = dict(gas=[dev(a, torch.float64) for a in O.gas_ic(g, 5)])
        phys = qj.Physics(rho_bar=1.2, four_pi_G=20.0, cs=0.3)
    a = qj.KDKSimulation(qj.Plan(16, dtype=dtype), phys, 1e-3, psi, **kw)
    a.step(3)
    ck.save_checkpoint(a, str(tmp_path), 3, params={"nx": 16, "dt": 1e-3})
    d = ck.load_checkpoint(str(tmp_path), 3)
    assert d["t"] == pytest.approx(3e-3)
    b = ck.restore_simulation(qj.Plan(16, dtype=dtype), phys, 1e-3, str(tmp_path), 3)
    a.step(2)
    b.step(2)
    sa, sb = a.state(), b.state()
    for k in sa:
        if isinstance(sa[k], torch.Tensor):
            err = ((sa[k] - sb[k]).abs().max() / sa[k].abs().max()).item()
            assert err < 1e-12, k
        else:
            assert sb[k] == pytest.approx(sa[k], rel=1e-12), k
