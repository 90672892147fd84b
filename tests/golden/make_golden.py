"""Writes the committed regression fixtures of tests/golden/ from the oracle.

The reference (Python/JAX) cannot run in this environment (no JAX), so
these vectors are FP64 outputs of the oracle on small seeded inputs, not
outputs of the reference itself.  They fix the oracle's results (the CPU
suite recomputes them and must agree) and give the GPU suite inputs and
expected outputs that need no SciPy at run time.

Run from the repository root:  python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402


def cases():
    """name -> (inputs dict, expected outputs dict)."""
    out = {}
    # plain KDK step (C1 in miniature): 16^3 c128, 3 steps
    g = O.Grid(16, 16, 16)
    ph = O.Physics(m_per_hbar=1.0, four_pi_G=1.0, rho_bar=1.0)
    psi = O.filtered_gaussian_psi(g, 101)
    p = psi
    for _ in range(3):
        p = O.compute_step_plain(p, 1e-3, g, ph)
    out["plain_16"] = ({"psi": psi, "dt": 1e-3, "steps": 3},
                       {"psi": p, "V": O.get_potential(np.abs(p) ** 2, g, ph)})
    # stars step: anisotropic 8x16x32 box, per-star masses, 2 steps
    g = O.Grid(8, 16, 32, 1.0, 2.0, 0.5)
    ph = O.Physics(rho_bar=1.1, four_pi_G=25.0)
    psi = O.filtered_gaussian_psi(g, 102, sigma_k=6 * np.pi)
    pos, vel, m_s = O.star_ic(g, 64, 1102, sigma_v=0.2)
    m = m_s * (0.5 + np.random.default_rng(7).random(64))
    s = (psi, pos, vel, 0.0, 0.0, np.inf)
    for _ in range(2):
        s = O.compute_step_stars(*s, 2e-3, m, g, ph)
    out["stars_8x16x32"] = (
        {"psi": psi, "pos": pos, "vel": vel, "mass": m, "dt": 2e-3, "steps": 2,
         "L": np.array([1.0, 2.0, 0.5]), "rho_bar": 1.1, "four_pi_G": 25.0},
        {"psi": s[0], "pos": s[1], "vel": s[2], "a_max": s[4], "dt_s": s[5]})
    # CIC indices and weights at faces, centres and their FP64 neighbours
    g = O.Grid(16, 16, 16)
    dx = 1.0 / 16
    base = np.concatenate([np.arange(17) * dx, (np.arange(16) + 0.5) * dx, [-0.3 * dx, 1.0]])
    xs = np.concatenate([base, np.nextafter(base, -1), np.nextafter(base, 2)])
    pos = np.stack([xs, np.roll(xs, 7), np.roll(xs, 19)], axis=1)
    i, ip1, wi, wip1 = O.get_cic_indices_and_weights(pos, g)
    out["cic_16"] = ({"pos": pos}, {"i": i, "ip1": ip1, "w_i": wi, "w_ip1": wip1})
    # one MUSCL update (16^3, random velocities)
    g = O.Grid(16, 16, 16)
    rho, _, _, _ = O.gas_ic(g, 103, sigma_k=6 * np.pi)
    rng = np.random.default_rng(8)
    vx, vy, vz = (0.1 * rng.standard_normal(g.shape) for _ in range(3))
    dx = 1.0 / 16
    r = O.solve_hydro(rho, vx, vy, vz, 2e-3, dx, dx ** 3, 0.3)
    out["hydro_16"] = ({"rho": rho, "vx": vx, "vy": vy, "vz": vz, "dt": 2e-3, "cs": 0.3},
                       {"rho": r[0], "vx": r[1], "vy": r[2], "vz": r[3]})
    return out


def main():
    for name, (inp, exp) in cases().items():
        arrays = {f"in_{k}": np.asarray(v) for k, v in inp.items()}
        arrays.update({f"out_{k}": np.asarray(v) for k, v in exp.items()})
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
        print(name, sum(a.nbytes for a in arrays.values()), "bytes")


if __name__ == "__main__":
    main()
