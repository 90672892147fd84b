"""qj_kdk_steps_noalloc against qj_kdk_steps (needs a GPU).

The no-allocation stepper runs the same kernels with its workspace taken
from the library's static per-GPU array instead of a caller buffer.  Two
calls in a row on one GPU, the second on a smaller grid (the array keeps the
first call's contents): a 128^3 complex64 particle-free state, whose steps
are deterministic and so must match qj_kdk_steps bit for bit, then a 32^3
complex128 stars state (its deposit adds with LDS float atomics in a
run-dependent order, so it matches to 1e-12).  A plan whose workspace
exceeds the array is rejected before any device work.
"""

import ctypes as C
import importlib

import numpy as np
import pytest
import torch

import oracle as O

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128


def dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x)).to("cuda", dtype)


def _run(qj, sim, nsteps, noalloc):
    _lib = importlib.import_module("quantum-jax_amd._lib")
    L = _lib.lib()
    with torch.cuda.device(sim.plan.device):
        if noalloc:
            rc = L.qj_kdk_steps_noalloc(sim.plan.handle, C.byref(sim._st), C.byref(sim._params),
                                        nsteps, sim.plan.stream())
        else:
            rc = L.qj_kdk_steps(sim.plan.handle, C.byref(sim._st), C.byref(sim._params), nsteps,
                                C.c_void_p(sim.plan.workspace().data_ptr()), sim.plan.stream())
    _lib.check(rc, "noalloc test")
    sim.t += nsteps * sim.dt
    return sim.state()


def test_noalloc_steps_match_workspace_steps(qj, cuda):
    _lib = importlib.import_module("quantum-jax_amd._lib")
    # call 1: 128^3 c64, no particles (a workspace of 34 MB of the 40 MiB)
    g = O.Grid(128, 128, 128)
    psi = dev(O.filtered_gaussian_psi(g, 51, sigma_k=16 * np.pi), C64)
    out = []
    for noalloc in (False, True):
        plan = qj.Plan(128, dtype=C64)
        assert _lib.lib().qj_workspace_size(plan.handle) <= _lib.NOALLOC_SCRATCH_BYTES
        sim = qj.KDKSimulation(plan, qj.Physics(four_pi_G=20.0), 1e-3, psi)
        st = _run(qj, sim, 3, noalloc)
        out.append((st["psi"].clone(), sim.V.clone(), sim.diag.clone()))
        del sim, plan
    for a, b in zip(*out):
        assert torch.equal(a, b)

    # call 2: 32^3 c128 with stars (smaller than call 1)
    g = O.Grid(32, 32, 32)
    psi = dev(O.filtered_gaussian_psi(g, 52, sigma_k=8 * np.pi), C128)
    pos, vel, m_s = O.star_ic(g, 2000, 1052, sigma_v=0.2)
    out = []
    for noalloc in (False, True):
        plan = qj.Plan(32, dtype=C128)
        sim = qj.KDKSimulation(plan, qj.Physics(rho_bar=1.1, four_pi_G=20.0), 1e-3, psi,
                               dev(pos, torch.float64), dev(vel, torch.float64), mass=m_s)
        st = _run(qj, sim, 2, noalloc)
        out.append((st["psi"].clone(), st["pos"].clone(), st["vel"].clone(), sim.V.clone()))
        del sim, plan
    rel = lambda a, b: ((a - b).norm() / b.norm()).item()
    for a, b in zip(*out):
        assert rel(a, b) < 1e-12


def test_noalloc_rejects_plans_beyond_its_workspace(qj, cuda):
    _lib = importlib.import_module("quantum-jax_amd._lib")
    plan = qj.Plan(256, 256, 128, dtype=C64)
    assert _lib.lib().qj_workspace_size(plan.handle) > _lib.NOALLOC_SCRATCH_BYTES
    st, prm = _lib.QjState(), _lib.QjParams(1e-3, 1.0, 1.0, 1.0, 0.0, 0.1, 0, 0)
    rc = _lib.lib().qj_kdk_steps_noalloc(plan.handle, C.byref(st), C.byref(prm), 1,
                                         plan.stream())
    assert rc == -2
    assert b"exceeds" in _lib.lib().qj_last_error()
