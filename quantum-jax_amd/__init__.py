"""MI355X-native (gfx950) split-step KDK Schroedinger-Poisson step.

A drop-in for the hot path of pmocz/quantum-jax: the potential kick, the
kinetic drift, the FFT Poisson solve, CIC star deposit/interpolation and the
MUSCL isothermal-gas update, with the reference's operator names, running on
hand-written HIP kernels through the C ABI of ``include/qj_api.h``.

Import with ``importlib.import_module("quantum-jax_amd")`` (the directory
name carries a hyphen).
"""

from .core import Physics, Plan  # noqa: F401
from .operators import (  # noqa: F401
    acceleration_grid, apply_grav_accel, bin_particles, bin_stars, do_accretion, drift,
    get_acceleration, get_cic_indices_and_weights, get_potential, interp, kick,
    solve_hydro,
)
from .simulation import BHSimulation, KDKSimulation, compute_step, compute_step_gas  # noqa: F401
from .ensemble import Ensemble  # noqa: F401
from . import accretion, checkpoint, cosmo, hdf5, history, ic  # noqa: F401
from ._lib import LIB_PATH, QjError, lib  # noqa: F401

__version__ = "0.1.0"
