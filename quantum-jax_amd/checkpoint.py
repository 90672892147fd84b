"""Snapshot I/O of the stepper state.

The reference saves its state pytree every chunk with orbax
(quantum-stars.py:162-165,449-456: psi, pos, vel, t, a_max, dt_s; the gas
run adds rho, vx, vy, vz) and writes its parameters to params.json.  orbax
is not available here; the same fields go to ``<directory>/<step>/state.npz``
(NumPy, read back with allow_pickle=False) and the parameters to
``<directory>/params.json``.  A black-hole run (quantum-bh.py:345-381) also
saves its accretion constants (G, c, sigma), and restores as a BHSimulation.
"""

from __future__ import annotations

import json
import os

import numpy as np
import torch

from . import _lib
from .core import Physics, Plan
from .simulation import BHSimulation, KDKSimulation


def save_checkpoint(sim: KDKSimulation, directory: str, step: int,
                    params: dict | None = None) -> str:
    """Write the end-of-step state (owed kicks applied) of sim."""
    st = sim.state()
    arrays = {"psi": st["psi"].cpu().numpy(), "t": np.float64(st["t"])}
    for k in ("pos", "vel", "rho", "vx", "vy", "vz"):
        if k in st:
            arrays[k] = st[k].cpu().numpy()
    if sim.np:
        arrays["a_max"] = np.float64(st["a_max"])
        arrays["dt_s"] = np.float64(st["dt_s"])
        arrays["mass"] = (sim.mass_t.cpu().numpy() if sim.mass_t is not None
                          else np.float64(sim.m_scalar))
    if isinstance(sim, BHSimulation):
        arrays["bh_G"] = np.float64(sim.G)
        arrays["bh_c"] = np.float64(sim.c)
        arrays["bh_sigma"] = np.float64(sim.sigma)
    path = os.path.join(directory, str(int(step)))
    os.makedirs(path, exist_ok=True)
    np.savez(os.path.join(path, "state.npz"), **arrays)
    if params is not None:
        with open(os.path.join(directory, "params.json"), "w") as fh:
            json.dump(params, fh, indent=2)
    return path


def load_checkpoint(directory: str, step: int) -> dict:
    """The saved fields as NumPy arrays (scalars as floats)."""
    with np.load(os.path.join(directory, str(int(step)), "state.npz"),
                 allow_pickle=False) as z:
        return {k: (z[k].item() if z[k].ndim == 0 else z[k]) for k in z.files}


def restore_simulation(plan: Plan, phys: Physics, dt: float, directory: str,
                       step: int) -> KDKSimulation:
    """A KDKSimulation at the saved state (a_max carried over); a
    BHSimulation with its saved accretion constants when the checkpoint is
    one of a black-hole run."""
    d = load_checkpoint(directory, step)
    dev = plan.device
    psi = torch.as_tensor(d["psi"]).to(dev, plan.complex_dtype)
    kw = {"t": float(d["t"])}
    if "pos" in d:
        m = d["mass"]
        kw.update(pos=torch.as_tensor(d["pos"]).to(dev), vel=torch.as_tensor(d["vel"]).to(dev),
                  mass=(torch.as_tensor(m).to(dev) if np.ndim(m) else float(m)))
    if "rho" in d:
        kw["gas"] = [torch.as_tensor(d[k]).to(dev, plan.real_dtype) for k in ("rho", "vx", "vy", "vz")]
    if "bh_G" in d:
        sim = BHSimulation(plan, phys, dt, psi, kw["pos"], kw["vel"],
                           torch.as_tensor(np.atleast_1d(d["mass"])).to(dev, torch.float64),
                           G=d["bh_G"], c=d["bh_c"], sigma=d["bh_sigma"], t=kw["t"])
    else:
        sim = KDKSimulation(plan, phys, dt, psi, **kw)
    if "a_max" in d:
        sim.diag[_lib.DIAG_AMAX] = float(d["a_max"])
    return sim
