"""Snapshot I/O of the stepper state.

The reference saves its state pytree every chunk with orbax
(quantum-stars.py:162-165,449-456: psi, pos, vel, t, a_max, dt_s; the gas
run adds rho, vx, vy, vz) and writes its parameters to params.json.  orbax
is not available here; the same fields go to ``<directory>/<step>/state.npz``
(NumPy, read back with allow_pickle=False) and the parameters to
``<directory>/params.json``.
"""

from __future__ import annotations

import json
import os

import numpy as np
import torch

from .core import Physics, Plan
from .simulation import KDKSimulation


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
    """A KDKSimulation at the saved state (a_max carried over)."""
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
    sim = KDKSimulation(plan, phys, dt, psi, **kw)
    if "a_max" in d:
        sim.diag[0] = float(d["a_max"])
    return sim
