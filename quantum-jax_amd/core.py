"""Plan (grid, box, precision) and physical parameters.

The reference keeps the grid, k-arrays and constants as module globals that
every operator reads (quantum-stars.py:72-157, quantum-fluid.py:53-139).  Here
they become an explicit ``Plan`` (immutable grid/box/precision, owning the
native plan with its twiddle and phase tables plus a reusable workspace) and
a ``Physics`` record of the scalar constants.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib


@dataclass(frozen=True)
class Physics:
    """Scalar constants of the step (dimensionless units).

    m_per_hbar  quantum-stars.py:109        four_pi_G  4*pi*G of :191
    rho_bar     quantum-stars.py:83         b_coeff    self-interaction, :117-125
    cs          quantum-fluid.py:71         dc_mode    0: reference k=0 rule
                                                        1: zero-mean potential
    flux        gas face flux: "rusanov" (quantum-fluid.py:238-264, the
                reference's) or "hll" (HLL with Davis signal speeds)
    """
    m_per_hbar: float = 1.0
    four_pi_G: float = 1.0
    rho_bar: float = 1.0
    b_coeff: float = 0.0
    cs: float = 0.1
    dc_mode: int = 0
    flux: str = "rusanov"

    @property
    def flux_kind(self) -> int:
        if self.flux not in _lib.FLUXES:
            raise ValueError(f"flux must be one of {sorted(_lib.FLUXES)}")
        return _lib.FLUXES[self.flux]


def _is_pow2(n: int) -> bool:
    return n >= 8 and n <= 512 and (n & (n - 1)) == 0


class Plan:
    """Grid (nx, ny, nz), box (Lx, Ly, Lz) and precision of a KDK solve.

    dtype is torch.complex64 (FP32 fields) or torch.complex128 (FP64 fields).
    Star positions/velocities/masses are always float64.
    """

    def __init__(self, nx: int, ny: int | None = None, nz: int | None = None,
                 L=(1.0, 1.0, 1.0), dtype=torch.complex64, device=None):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        if not all(_is_pow2(n) for n in (nx, ny, nz)):
            raise ValueError("grid sizes must be powers of two in [8, 512]")
        if isinstance(L, (int, float)):
            L = (float(L),) * 3
        if dtype not in (torch.complex64, torch.complex128):
            raise ValueError("dtype must be torch.complex64 or torch.complex128")
        if not torch.cuda.is_available():
            raise RuntimeError("a HIP device is required (no CPU path)")
        dev = torch.device(device) if device is not None else torch.device("cuda")
        if dev.type != "cuda":
            raise ValueError("the plan needs a HIP device (torch 'cuda' device)")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.nx, self.ny, self.nz = nx, ny, nz
        self.Lx, self.Ly, self.Lz = (float(v) for v in L)
        self.complex_dtype = dtype
        self.real_dtype = torch.float32 if dtype == torch.complex64 else torch.float64
        self._handle = C.c_void_p()
        lib = _lib.lib()
        with torch.cuda.device(self.device):
            _lib.check(lib.qj_plan_create(C.byref(self._handle), nx, ny, nz, self.Lx,
                                          self.Ly, self.Lz,
                                          _lib.QJ_C64 if dtype == torch.complex64
                                          else _lib.QJ_C128), "qj_plan_create")
        self._work = None

    # -- geometry --------------------------------------------------------
    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)

    @property
    def ncell(self):
        return self.nx * self.ny * self.nz

    @property
    def dx(self):
        return self.Lx / self.nx

    @property
    def dy(self):
        return self.Ly / self.ny

    @property
    def dz(self):
        return self.Lz / self.nz

    @property
    def vol(self):
        return self.dx * self.dy * self.dz

    # -- native handles ----------------------------------------------------
    @property
    def handle(self):
        return self._handle

    def workspace(self) -> torch.Tensor:
        if self._work is None:
            n = int(_lib.lib().qj_workspace_size(self._handle))
            self._work = torch.empty(n, dtype=torch.uint8, device=self.device)
        return self._work

    def set_option(self, option: int, value: int):
        """Schedule option (e.g. _lib.OPT_FUSE_X_SOURCE); results are unchanged."""
        _lib.check(_lib.lib().qj_plan_set_option(self._handle, int(option), int(value)),
                   "qj_plan_set_option")

    def get_option(self, option: int) -> int:
        v = C.c_int()
        _lib.check(_lib.lib().qj_plan_get_option(self._handle, int(option), C.byref(v)),
                   "qj_plan_get_option")
        return v.value

    def stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def close(self):
        if getattr(self, "_handle", None) and self._handle.value:
            try:
                _lib.lib().qj_plan_destroy(self._handle)
            finally:
                self._handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- tensor checks -------------------------------------------------------
    def field(self, t: torch.Tensor, name: str, complex_: bool = False, planes: int = 0):
        want = self.complex_dtype if complex_ else self.real_dtype
        shape = ((planes,) if planes else ()) + self.shape
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch tensor")
        if t.device != self.device:
            raise ValueError(f"{name} must live on {self.device}, got {t.device}")
        if t.dtype != want:
            raise ValueError(f"{name} must be {want}, got {t.dtype}")
        if tuple(t.shape) != shape:
            raise ValueError(f"{name} must have shape {shape}, got {tuple(t.shape)}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        return t

    def particles(self, t: torch.Tensor, name: str, cols: int = 3):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch tensor")
        if t.device != self.device or t.dtype != torch.float64:
            raise ValueError(f"{name} must be float64 on {self.device}")
        if cols and (t.dim() != 2 or t.shape[1] != cols):
            raise ValueError(f"{name} must have shape (np, {cols})")
        if not cols and t.dim() != 1:
            raise ValueError(f"{name} must have shape (np,)")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
        return t

    def new_diag(self) -> torch.Tensor:
        d = torch.zeros(_lib.DIAG_COUNT, dtype=torch.float64, device=self.device)
        d[_lib.DIAG_DTS] = float("inf")
        return d


def ptr(t) -> C.c_void_p:
    return C.c_void_p(0 if t is None else t.data_ptr())
