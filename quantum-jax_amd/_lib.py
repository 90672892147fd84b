"""ctypes binding of libqjamd.so (the C ABI declared in include/qj_api.h).

The library is required: there is no fallback.  If it has not been built,
``lib()`` raises with instructions instead of silently running anything else.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqjamd.so")

QJ_C64 = 0
QJ_C128 = 1
DIAG_AMAX, DIAG_DTS, DIAG_MASS, DIAG_NAN, DIAG_COUNT = 0, 1, 2, 3, 8
# reals per cell of the acceleration grid (a_x, a_y, a_z); the state's grid
# holds V (one real per cell) and then this grid
ACC_COMPONENTS = 3
FLUX_RUSANOV, FLUX_HLL = 0, 1
FLUXES = {"rusanov": FLUX_RUSANOV, "hll": FLUX_HLL}
OPT_FUSE_X_SOURCE = 0
OPT_BINNED_DEPOSIT = 1
OPT_DETERMINISTIC_DEPOSIT = 2
OPT_WIDE_SWEEPS = 3
OPT_Z_SWEEP_BLOCKS = 4
OPT_SIDE_STREAM = 5
OPT_STAR_GATHER = 6
OPT_GATHER_ROWS = 7
OPT_FUSED_DEPOSIT = 8
OPT_COUNT = 9
ENS_EAGER, ENS_BATCHED, ENS_GRAPH, ENS_GRAPH_FORK = 0, 1, 2, 3
# bytes of the static workspace of qj_kdk_steps_noalloc (one per GPU)
NOALLOC_SCRATCH_BYTES = 40 * 1024 * 1024

vp = C.c_void_p
dp = C.POINTER(C.c_double)
i64 = C.c_longlong
f64 = C.c_double
i32 = C.c_int


class QjState(C.Structure):
    """Mirror of ``qj_state``."""
    _fields_ = [
        ("psi", vp), ("rho_s", vp), ("gas", vp * 4), ("gas_tmp", vp * 4),
        ("pos", vp), ("vel", vp), ("mass", vp), ("m_scalar", f64), ("np", i64),
        ("grid", vp), ("brho", vp), ("diag", vp),
        ("grid_valid", i32), ("gas_swapped", i32), ("pending", f64), ("pending_star", f64),
        ("star_scratch", vp), ("v_ext", vp), ("rho_valid", i32),
    ]


class QjParams(C.Structure):
    """Mirror of ``qj_params``."""
    _fields_ = [
        ("dt", f64), ("m_per_hbar", f64), ("four_pi_G", f64), ("rho_bar", f64),
        ("b_coeff", f64), ("cs", f64), ("dc_mode", i32), ("flux", i32),
    ]


class QjTapeStep(C.Structure):
    """Mirror of ``qj_tape_step``."""
    _fields_ = [("psi0", vp), ("psi2", vp), ("grid1", vp), ("x0", vp), ("x1", vp),
                ("bins", vp)]


class QjAdjoint(C.Structure):
    """Mirror of ``qj_adjoint``."""
    _fields_ = [("psi_bar", vp), ("x_bar", vp), ("v_bar", vp), ("m_bar", vp), ("scratch", vp)]


# name -> (restype, argtypes); every entry point returns int status unless noted
SIGNATURES = {
    "qj_plan_create": (i32, [C.POINTER(vp), i32, i32, i32, f64, f64, f64, i32]),
    "qj_plan_destroy": (i32, [vp]),
    "qj_workspace_size": (C.c_size_t, [vp]),
    "qj_last_error": (C.c_char_p, []),
    "qj_version": (i32, []),
    "qj_get_potential": (i32, [vp, vp, f64, f64, f64, i32, vp, vp, vp]),
    "qj_acceleration_grid": (i32, [vp, vp, f64, f64, i32, vp, vp, vp, vp, vp]),
    "qj_acceleration_grid_lp": (i32, [vp, vp, f64, f64, f64, vp, vp, vp, vp]),
    "qj_bondi_sink": (i32, [vp, vp, vp, vp, vp, vp, f64, f64, f64, f64, f64, f64, f64, f64, vp,
                            vp]),
    "qj_kick": (i32, [vp, vp, vp, f64, vp]),
    "qj_drift": (i32, [vp, vp, f64, f64, vp]),
    "qj_cic_indices": (i32, [vp, vp, i64, vp, vp, vp, vp, vp]),
    "qj_bin_particles": (i32, [vp, vp, vp, f64, i64, vp, i32, vp]),
    "qj_interp": (i32, [vp, vp, i64, vp, vp, vp]),
    "qj_get_acceleration": (i32, [vp, vp, i64, vp, f64, f64, i32, vp, vp, vp, vp, vp]),
    "qj_apply_grav_accel": (i32, [vp, vp, vp, vp, vp, f64, f64, f64, i32, vp, vp, vp, vp]),
    "qj_solve_hydro": (i32, [vp, vp, vp, vp, vp, f64, f64, i32, vp, vp, vp, vp, vp]),
    "qj_gas_cfl_dt": (i32, [vp, vp, vp, f64, f64, vp, vp]),
    "qj_kdk_steps_cfl": (i32, [vp, vp, vp, i32, f64, f64, dp, vp, vp]),
    "qj_kdk_steps": (i32, [vp, C.POINTER(QjState), C.POINTER(QjParams), i32, vp, vp]),
    "qj_kdk_finalize": (i32, [vp, C.POINTER(QjState), C.POINTER(QjParams), vp]),
    "qj_star_density": (i32, [vp, C.POINTER(QjState), vp]),
    "qj_kdk_steps_noalloc": (i32, [vp, C.POINTER(QjState), C.POINTER(QjParams), i32, vp]),
    "qj_kdk_steps_sched": (i32, [vp, C.POINTER(QjState), C.POINTER(QjParams), i32, dp, dp, vp, vp]),
    "qj_adjoint_scratch_size": (C.c_size_t, [vp]),
    "qj_star_scratch_size": (C.c_size_t, [vp, i64]),
    "qj_bh_accretion": (i32, [vp, vp, vp, vp, vp, i64, f64, f64, f64, f64, f64, vp]),
    "qj_kdk_record": (i32, [vp, C.POINTER(QjState), C.POINTER(QjParams), i32,
                            C.POINTER(QjTapeStep), vp, vp, vp]),
    "qj_kdk_adjoint": (i32, [vp, C.POINTER(QjState), C.POINTER(QjParams), i32,
                             C.POINTER(QjTapeStep), vp, C.POINTER(QjAdjoint), vp, vp]),
    "qj_deposit_adjoint": (i32, [vp, vp, vp, f64, i64, vp, vp, vp, vp]),
    "qj_plan_set_option": (i32, [vp, i32, i32]),
    "qj_plan_get_option": (i32, [vp, i32, C.POINTER(C.c_int)]),
    "qj_ensemble_create": (i32, [C.POINTER(vp), i32, C.POINTER(vp), C.POINTER(C.POINTER(QjState)),
                                 C.POINTER(QjParams), C.POINTER(vp)]),
    "qj_ensemble_destroy": (i32, [vp]),
    "qj_ensemble_set_mode": (i32, [vp, i32]),
    "qj_ensemble_steps": (i32, [vp, i32, vp]),
    "qj_ensemble_finalize": (i32, [vp, vp]),
    "qj_ensemble_record": (i32, [vp, i32, C.POINTER(C.POINTER(QjTapeStep)), C.POINTER(vp), vp]),
    "qj_ensemble_adjoint": (i32, [vp, i32, C.POINTER(C.POINTER(QjTapeStep)), C.POINTER(vp),
                                  C.POINTER(QjAdjoint), vp]),
    "qj_ensemble_group_count": (i32, [vp]),
    "qj_ensemble_group_info": (i32, [vp, i32, C.POINTER(C.c_int), i32, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]),
    "qj_stage_count": (i32, []),
    "qj_stage_name": (C.c_char_p, [i32]),
    "qj_timing_enable": (i32, [vp, i32]),
    "qj_timing_read": (i32, [vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong), i32]),
}

_lock = threading.Lock()
_lib = None


class QjError(RuntimeError):
    pass


def lib():
    """Load libqjamd.so (once).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise QjError(
                f"{LIB_PATH} is missing: build the gfx950 library first "
                "(python quantum-jax_amd/build.py or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().qj_last_error().decode(errors="replace")
        kind = {-1: "bad argument", -2: "unsupported", -3: "HIP error"}.get(rc, str(rc))
        raise QjError(f"{what}: {kind}: {msg}")
