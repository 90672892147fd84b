"""FP64 NumPy/SciPy restatement of the reference KDK Schroedinger-Poisson path.

TEST INFRASTRUCTURE ONLY (see the package docstring): the checker for the HIP
path and the timed "CPU restatement of the reference path".

Every operator follows the reference scripts in their own order (unmerged
half-kicks, the reference FFT count), with the reference's module globals made
explicit through ``Grid`` and ``Physics``.  Arrays are C-ordered ``(nx,ny,nz)``
with ``indexing="ij"`` (z contiguous), as in ``quantum-stars.py:139,145``.

Parity status per function (pins live in ``tests/test_oracle.py``):
  make_grid / kvec            pinned (FFT-order identity, Nyquist sign)
  get_potential               pinned (single-mode closed form, k=0 rule, Laplacian)
  kick                        pinned (|psi| invariant, constant-V global phase)
  drift                       pinned (plane-wave closed form, unitarity, reversal)
  get_cic_indices_and_weights pinned (exact rational floor away from faces,
                              centre/face cases, wrap cases, brute force)
  bin_particles / bin_stars   pinned (mass conservation, centre deposit,
                              deposit/interp transpose identity)
  get_acceleration            pinned (single-mode analytic gradient, constant
                              and linear-field interpolation)
  compute_step_stars/_plain   pinned (time reversibility, 2nd order in dt,
                              mass conservation)
  hydro (solve_hydro & co.)   pinned (uniform state, conservation, minmod
                              identity, sound speed, 1-D Riemann symmetry,
                              exact isothermal Riemann solution)
  cfl_dt                      pinned (Courant number of a uniform flow equals
                              cfl; the dt_max cap)
  get_flux_hll                pinned (consistency F(U,U) = F(U), supersonic
                              upwinding, the HLL integral-average form,
                              exact isothermal Riemann solution)
  compute_step_gas            pinned (mass conservation, reversibility of the
                              spectral part, uniform-gas reduction to plain)
  lp_factor /                 pinned (single-mode closed form: the filtered
  acceleration_grid_lp        gradient is the unfiltered one times the factor;
                              k = 0 passes; large k_cut = unfiltered)
  sample_cic                  pinned (constant and periodic-linear fields exact;
                              equal to interp_grid's CIC)
  bh_bondi_step_ambient       pinned (uniform gas: rho_inf exact, Bondi rate
                              closed form, removed mass = dM_cap, removal
                              profile W rho inside r_acc only; BHL rate with a
                              uniform flow; the M_avail cap)
  seed_mass_from_time         pinned (limits of the tanh ramp, the step form)
Unpinned: the reference's FP32 rounding trajectory (needs a JAX import).
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np
import scipy.fft as sfft

__all__ = [
    "Grid", "Physics", "kvec", "make_grid", "cell_centres", "set_workers", "CORNERS",
    "fftn", "ifftn",
    "get_potential", "kick", "drift",
    "get_cic_indices_and_weights", "bin_particles", "bin_stars",
    "interp_grid", "acceleration_grid", "get_acceleration",
    "compute_step_stars", "compute_step_plain",
    "get_conserved", "get_primitive", "get_gradient", "extrap_to_face",
    "apply_fluxes", "get_flux", "get_flux_hll", "slope_limiter", "solve_hydro", "cfl_dt",
    "apply_grav_accel", "compute_step_gas",
    "mass_total", "energy_total",
    "filtered_gaussian_psi", "filtered_gaussian_field", "star_ic", "gas_ic",
    "compute_dt_cosmo", "compute_step_cosmo", "external_point_mass", "compute_step_tidal",
    "do_accretion", "compute_step_bh",
    "lp_factor", "acceleration_grid_lp", "sample_cic", "bh_bondi_step_ambient",
    "seed_mass_from_time",
]

_WORKERS = 1


def set_workers(n: int) -> None:
    """Number of threads scipy.fft may use (1 = scalar baseline)."""
    global _WORKERS
    _WORKERS = max(1, int(n))


def fftn(a):
    return sfft.fftn(a, workers=_WORKERS)


def ifftn(a):
    return sfft.ifftn(a, workers=_WORKERS)


# --------------------------------------------------------------------------
# Mesh (quantum-stars.py:131-149; quantum-fluid.py:115-131)
# --------------------------------------------------------------------------

@dataclass(frozen=True)
class Grid:
    nx: int
    ny: int
    nz: int
    Lx: float = 1.0
    Ly: float = 1.0
    Lz: float = 1.0

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)

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
        # quantum-stars.py:135  vol = dx * dy * dz
        return self.dx * self.dy * self.dz

    @property
    def ncell(self):
        return self.nx * self.ny * self.nz


@dataclass(frozen=True)
class Physics:
    """Explicit form of the reference's module constants (dimensionless)."""
    m_per_hbar: float = 1.0      # quantum-stars.py:109
    four_pi_G: float = 1.0       # 4*pi*G, quantum-stars.py:191
    rho_bar: float = 1.0         # quantum-stars.py:83
    b_coeff: float = 0.0         # quantum-stars.py:117-125 (self-interaction)
    cs: float = 0.1              # quantum-fluid.py:71 (sound speed)
    flux: str = "rusanov"        # gas face flux: "rusanov" (quantum-fluid.py:238-264) or "hll"


def kvec(n: int, L: float) -> np.ndarray:
    """1-D wavenumbers in FFT order, Nyquist bin negative.

    quantum-stars.py:142-148: ``2*pi/L * arange(-n/2, n/2)`` then ifftshift.
    """
    return np.fft.ifftshift(2.0 * np.pi / L * np.arange(-n / 2, n / 2))


@dataclass
class _Mesh:
    g: Grid
    kx: np.ndarray = field(init=False)
    ky: np.ndarray = field(init=False)
    kz: np.ndarray = field(init=False)
    k_sq: np.ndarray = field(init=False)

    def __post_init__(self):
        g = self.g
        kx = kvec(g.nx, g.Lx)[:, None, None]
        ky = kvec(g.ny, g.Ly)[None, :, None]
        kz = kvec(g.nz, g.Lz)[None, None, :]
        self.kx, self.ky, self.kz = kx, ky, kz
        # quantum-stars.py:149
        self.k_sq = kx ** 2 + ky ** 2 + kz ** 2


_MESH_CACHE: dict = {}


def make_grid(g: Grid) -> _Mesh:
    m = _MESH_CACHE.get(g)
    if m is None:
        m = _Mesh(g)
        _MESH_CACHE[g] = m
    return m


def cell_centres(g: Grid):
    """quantum-stars.py:136-139 (linspace of cell centres, indexing='ij')."""
    x = np.linspace(0.5 * g.dx, g.Lx - 0.5 * g.dx, g.nx)
    y = np.linspace(0.5 * g.dy, g.Ly - 0.5 * g.dy, g.ny)
    z = np.linspace(0.5 * g.dz, g.Lz - 0.5 * g.dz, g.nz)
    return np.meshgrid(x, y, z, indexing="ij")


# --------------------------------------------------------------------------
# Gravity (quantum-stars.py:189-193; quantum-fluid.py:171-175)
# --------------------------------------------------------------------------

def _v_hat(rho, g: Grid, ph: Physics):
    m = make_grid(g)
    k_sq = m.k_sq
    # quantum-stars.py:191 -- the k=0 mode is divided by 1, not zeroed
    return -fftn(ph.four_pi_G * (rho - ph.rho_bar)) / (k_sq + (k_sq == 0))


def get_potential(rho, g: Grid, ph: Physics):
    """quantum-stars.py:189-193."""
    V = np.real(ifftn(_v_hat(rho, g, ph)))
    return V + ph.b_coeff * rho if ph.b_coeff != 0.0 else V


def kick(psi, V, coef):
    """psi <- exp(-1j*coef*V)*psi, coef = m_per_hbar*dt/2 (quantum-stars.py:301)."""
    return np.exp(-1.0j * coef * V) * psi


def drift(psi, dt, g: Grid, ph: Physics):
    """Kinetic drift, quantum-stars.py:307-309."""
    m = make_grid(g)
    psi_hat = fftn(psi)
    psi_hat = np.exp(dt * (-1.0j * m.k_sq / ph.m_per_hbar / 2.0)) * psi_hat
    return ifftn(psi_hat)


# --------------------------------------------------------------------------
# Cloud-in-cell (quantum-stars.py:200-288; quantum-bh.py:223-258)
# --------------------------------------------------------------------------

def get_cic_indices_and_weights(pos, g: Grid):
    """quantum-stars.py:200-209, evaluated in IEEE FP64 (no FMA contraction)."""
    dxs = np.array([g.dx, g.dy, g.dz])
    i = np.floor((pos - 0.5 * dxs) / dxs)
    ip1 = i + 1.0
    weight_i = ((ip1 + 0.5) * dxs - pos) / dxs
    weight_ip1 = (pos - (i + 0.5) * dxs) / dxs
    n = np.array([g.nx, g.ny, g.nz])
    i = np.mod(i, n).astype(np.int64)
    ip1 = np.mod(ip1, n).astype(np.int64)
    return i, ip1, weight_i, weight_ip1


# Corner order of quantum-stars.py:220-243 / 264-287: (use_ip1_x, _y, _z).
CORNERS = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
           (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1))


def _corner_tables(pos, g: Grid):
    i, ip1, w_i, w_ip1 = get_cic_indices_and_weights(pos, g)
    idx = np.empty((pos.shape[0], 8), dtype=np.int64)
    w = np.empty((pos.shape[0], 8))
    for c, (sx, sy, sz) in enumerate(CORNERS):
        ix = ip1[:, 0] if sx else i[:, 0]
        iy = ip1[:, 1] if sy else i[:, 1]
        iz = ip1[:, 2] if sz else i[:, 2]
        wx = w_ip1[:, 0] if sx else w_i[:, 0]
        wy = w_ip1[:, 1] if sy else w_i[:, 1]
        wz = w_ip1[:, 2] if sz else w_i[:, 2]
        idx[:, c] = (ix * g.ny + iy) * g.nz + iz
        w[:, c] = wx * wy * wz
    return idx, w


def bin_particles(pos, m, g: Grid):
    """CIC mass deposit; quantum-bh.py:223-258 (per-particle m), star-major
    with the 8 corners in reference order, accumulated sequentially."""
    idx, w = _corner_tables(pos, g)
    m = np.broadcast_to(np.asarray(m, dtype=np.float64), (pos.shape[0],))
    fac = m / g.vol                                    # quantum-bh.py:230
    contrib = w * fac[:, None]
    rho = np.bincount(idx.ravel(), weights=contrib.ravel(), minlength=g.ncell)
    return rho.reshape(g.shape)


def bin_stars(pos, m_s, g: Grid):
    """quantum-stars.py:212-247 (scalar star mass m_s)."""
    return bin_particles(pos, m_s, g)


def interp_grid(fields, pos, g: Grid):
    """8-corner CIC gather of a stack of grid fields, quantum-stars.py:263-287.

    ``fields`` has shape (C, nx, ny, nz); returns (Np, C).
    """
    idx, w = _corner_tables(pos, g)
    flat = fields.reshape(fields.shape[0], -1)
    out = np.zeros((pos.shape[0], fields.shape[0]))
    for c in range(8):
        out += w[:, c:c + 1] * flat[:, idx[:, c]].T
    return out


def acceleration_grid(rho, g: Grid, ph: Physics):
    """a_d = -Re IFFT(1j*k_d*V_hat); quantum-stars.py:254-260."""
    m = make_grid(g)
    V_hat = _v_hat(rho, g, ph)
    ax = -np.real(ifftn(1.0j * m.kx * V_hat))
    ay = -np.real(ifftn(1.0j * m.ky * V_hat))
    az = -np.real(ifftn(1.0j * m.kz * V_hat))
    return np.stack((ax, ay, az))


def get_acceleration(pos, rho, g: Grid, ph: Physics):
    """quantum-stars.py:250-288 -> (acc (Np,3), a_max)."""
    a_grid = acceleration_grid(rho, g, ph)
    a_max = np.max(np.abs(a_grid))
    return interp_grid(a_grid, pos, g), a_max


# --------------------------------------------------------------------------
# KDK steps
# --------------------------------------------------------------------------

def compute_step_stars(psi, pos, vel, t, a_max, dt_s, dt, m_s, g: Grid,
                       ph: Physics):
    """quantum-stars.py:295-330 (m_s scalar or per-particle array)."""
    L = np.array([g.Lx, g.Ly, g.Lz])
    # (1/2) kick
    rho_s = bin_particles(pos, m_s, g)
    rho_tot = np.abs(psi) ** 2 + rho_s
    V = get_potential(rho_tot, g, ph)
    psi = kick(psi, V, ph.m_per_hbar * dt / 2.0)
    acc, a_max1 = get_acceleration(pos, rho_tot, g, ph)
    vel = vel + acc * dt / 2.0
    # drift
    psi = drift(psi, dt, g, ph)
    pos = pos + vel * dt
    pos = np.mod(pos, L)
    # (1/2) kick
    rho_s = bin_particles(pos, m_s, g)
    rho_tot = np.abs(psi) ** 2 + rho_s
    V = get_potential(rho_tot, g, ph)
    psi = kick(psi, V, ph.m_per_hbar * dt / 2.0)
    acc, a_max2 = get_acceleration(pos, rho_tot, g, ph)
    vel = vel + acc * dt / 2.0
    t = t + dt
    a_max = max(a_max, a_max1, a_max2)
    # quantum-stars.py:328 -- dx for every component (reference quirk)
    dt_s = np.min(np.abs(g.dx / vel)) if vel.size else np.inf
    return psi, pos, vel, t, a_max, dt_s


def compute_step_plain(psi, dt, g: Grid, ph: Physics):
    """The stars step with no particles (quantum-stars.py:295-330, n_s = 0)."""
    V = get_potential(np.abs(psi) ** 2, g, ph)
    psi = kick(psi, V, ph.m_per_hbar * dt / 2.0)
    psi = drift(psi, dt, g, ph)
    V = get_potential(np.abs(psi) ** 2, g, ph)
    return kick(psi, V, ph.m_per_hbar * dt / 2.0)


# --------------------------------------------------------------------------
# Isothermal gas (quantum-fluid.py:182-438)
# --------------------------------------------------------------------------

def get_conserved(rho, vx, vy, vz, vol):
    """quantum-fluid.py:182-188."""
    return rho * vol, rho * vx * vol, rho * vy * vol, rho * vz * vol


def get_primitive(Mass, Momx, Momy, Momz, vol):
    """quantum-fluid.py:191-197."""
    return Mass / vol, Momx / Mass, Momy / Mass, Momz / Mass


def get_gradient(f, dx):
    """quantum-fluid.py:200-205 (central differences, periodic)."""
    return tuple((np.roll(f, -1, axis=a) - np.roll(f, 1, axis=a)) / (2.0 * dx)
                 for a in range(3))


def extrap_to_face(f, f_dx, f_dy, f_dz, dx):
    """quantum-fluid.py:208-221."""
    out = []
    for a, fd in enumerate((f_dx, f_dy, f_dz)):
        fL = f + 0.5 * fd * dx
        fR = np.roll(f - 0.5 * fd * dx, -1, axis=a)
        out += [fL, fR]
    return tuple(out)


def apply_fluxes(F, flux_F_X, flux_F_Y, flux_F_Z, dx, dt):
    """quantum-fluid.py:224-235."""
    fac = dt * dx * dx
    F = F - fac * flux_F_X
    F = F + fac * np.roll(flux_F_X, 1, axis=0)
    F = F - fac * flux_F_Y
    F = F + fac * np.roll(flux_F_Y, 1, axis=1)
    F = F - fac * flux_F_Z
    F = F + fac * np.roll(flux_F_Z, 1, axis=2)
    return F


def get_flux(rho_L, vx_L, vy_L, vz_L, rho_R, vx_R, vy_R, vz_R, cs):
    """Rusanov flux of the averaged state, quantum-fluid.py:238-264."""
    rho_star = 0.5 * (rho_L + rho_R)
    momx_star = 0.5 * (rho_L * vx_L + rho_R * vx_R)
    momy_star = 0.5 * (rho_L * vy_L + rho_R * vy_R)
    momz_star = 0.5 * (rho_L * vz_L + rho_R * vz_R)
    P_star = rho_star * cs * cs
    flux_Mass = momx_star
    flux_Momx = momx_star ** 2 / rho_star + P_star
    flux_Momy = momx_star * momy_star / rho_star
    flux_Momz = momx_star * momz_star / rho_star
    C = np.maximum(cs + np.abs(vx_L), cs + np.abs(vx_R))
    flux_Mass = flux_Mass - C * 0.5 * (rho_R - rho_L)
    flux_Momx = flux_Momx - C * 0.5 * (rho_R * vx_R - rho_L * vx_L)
    flux_Momy = flux_Momy - C * 0.5 * (rho_R * vy_R - rho_L * vy_L)
    flux_Momz = flux_Momz - C * 0.5 * (rho_R * vz_R - rho_L * vz_L)
    return flux_Mass, flux_Momx, flux_Momy, flux_Momz


def get_flux_hll(rho_L, vx_L, vy_L, vz_L, rho_R, vx_R, vy_R, vz_R, cs):
    """HLL flux of the isothermal system (not in the reference, whose flux is
    get_flux above; the benchmark description asks for HLL).

    Harten, Lax and van Leer (SIAM Rev. 25, 1983) with Davis's signal-speed
    estimates S_L = min(u_L, u_R) - cs, S_R = max(u_L, u_R) + cs:
    F = F_L if S_L >= 0, F_R if S_R <= 0, else
    (S_R F_L - S_L F_R + S_L S_R (U_R - U_L)) / (S_R - S_L),
    with U = (rho, rho u, rho v, rho w) and
    F(U) = (rho u, rho u^2 + rho cs^2, rho u v, rho u w); u is face-normal.
    """
    c2 = cs * cs
    mL, mR = rho_L * vx_L, rho_R * vx_R
    FL = (mL, mL * vx_L + rho_L * c2, mL * vy_L, mL * vz_L)
    FR = (mR, mR * vx_R + rho_R * c2, mR * vy_R, mR * vz_R)
    UL = (rho_L, mL, rho_L * vy_L, rho_L * vz_L)
    UR = (rho_R, mR, rho_R * vy_R, rho_R * vz_R)
    SL = np.minimum(vx_L, vx_R) - cs
    SR = np.maximum(vx_L, vx_R) + cs
    out = []
    for fl, fr, ul, ur in zip(FL, FR, UL, UR):
        mid = (SR * fl - SL * fr + SL * SR * (ur - ul)) / (SR - SL)
        out.append(np.where(SL >= 0, fl, np.where(SR <= 0, fr, mid)))
    return tuple(out)


def slope_limiter(f, dx, f_dx, f_dy, f_dz):
    """Ratio-form minmod of quantum-fluid.py:352-407 (eps = 1e-12)."""
    eps = 1.0e-12
    out = []
    for a, fd in enumerate((f_dx, f_dy, f_dz)):
        denom = np.where(fd > 0, fd + eps, np.where(fd < 0, fd - eps, eps))
        num = (f - np.roll(f, 1, axis=a)) / dx
        lim = np.maximum(0.0, np.minimum(1.0, num / denom))
        fd = lim * fd
        num = -(f - np.roll(f, -1, axis=a)) / dx
        lim = np.maximum(0.0, np.minimum(1.0, num / denom))
        fd = lim * fd
        out.append(fd)
    return tuple(out)


def solve_hydro(rho, vx, vy, vz, dt, dx, vol, cs, flux="rusanov"):
    """MUSCL-Hancock isothermal update, quantum-fluid.py:281-349; flux "hll"
    replaces the face flux by get_flux_hll (same rotated arguments)."""
    get_flux_ = {"rusanov": get_flux, "hll": get_flux_hll}[flux]
    rho_dx, rho_dy, rho_dz = get_gradient(rho, dx)
    vx_dx, vx_dy, vx_dz = get_gradient(vx, dx)
    vy_dx, vy_dy, vy_dz = get_gradient(vy, dx)
    vz_dx, vz_dy, vz_dz = get_gradient(vz, dx)

    rho_dx, rho_dy, rho_dz = slope_limiter(rho, dx, rho_dx, rho_dy, rho_dz)
    vx_dx, vx_dy, vx_dz = slope_limiter(vx, dx, vx_dx, vx_dy, vx_dz)
    vy_dx, vy_dy, vy_dz = slope_limiter(vy, dx, vy_dx, vy_dy, vy_dz)
    vz_dx, vz_dy, vz_dz = slope_limiter(vz, dx, vz_dx, vz_dy, vz_dz)

    rho_prime = rho - 0.5 * dt * (vx * rho_dx + rho * vx_dx + vy * rho_dy
                                  + rho * vy_dy + vz * rho_dz + rho * vz_dz)
    vx_prime = vx - 0.5 * dt * (vx * vx_dx + vy * vx_dy + vz * vx_dz
                                + (1.0 / rho) * rho_dx * cs * cs)
    vy_prime = vy - 0.5 * dt * (vx * vy_dx + vy * vy_dy + vz * vy_dz
                                + (1.0 / rho) * rho_dy * cs * cs)
    vz_prime = vz - 0.5 * dt * (vx * vz_dx + vy * vz_dy + vz * vz_dz
                                + (1.0 / rho) * rho_dz * cs * cs)

    rho_XL, rho_XR, rho_YL, rho_YR, rho_ZL, rho_ZR = extrap_to_face(
        rho_prime, rho_dx, rho_dy, rho_dz, dx)
    vx_XL, vx_XR, vx_YL, vx_YR, vx_ZL, vx_ZR = extrap_to_face(
        vx_prime, vx_dx, vx_dy, vx_dz, dx)
    vy_XL, vy_XR, vy_YL, vy_YR, vy_ZL, vy_ZR = extrap_to_face(
        vy_prime, vy_dx, vy_dy, vy_dz, dx)
    vz_XL, vz_XR, vz_YL, vz_YR, vz_ZL, vz_ZR = extrap_to_face(
        vz_prime, vz_dx, vz_dy, vz_dz, dx)

    fMX, fPxX, fPyX, fPzX = get_flux_(rho_XL, vx_XL, vy_XL, vz_XL,
                                      rho_XR, vx_XR, vy_XR, vz_XR, cs)
    # quantum-fluid.py:331-336: rotated argument order for Y and Z
    fMY, fPyY, fPzY, fPxY = get_flux_(rho_YL, vy_YL, vz_YL, vx_YL,
                                      rho_YR, vy_YR, vz_YR, vx_YR, cs)
    fMZ, fPzZ, fPxZ, fPyZ = get_flux_(rho_ZL, vz_ZL, vx_ZL, vy_ZL,
                                      rho_ZR, vz_ZR, vx_ZR, vy_ZR, cs)

    Mass, Momx, Momy, Momz = get_conserved(rho, vx, vy, vz, vol)
    Mass = apply_fluxes(Mass, fMX, fMY, fMZ, dx, dt)
    Momx = apply_fluxes(Momx, fPxX, fPxY, fPxZ, dx, dt)
    Momy = apply_fluxes(Momy, fPyX, fPyY, fPyZ, dx, dt)
    Momz = apply_fluxes(Momz, fPzX, fPzY, fPzZ, dx, dt)
    return get_primitive(Mass, Momx, Momy, Momz, vol)


def cfl_dt(vx, vy, vz, cs, dx, cfl, dt_max):
    """Hydro CFL cap of quantum-fluid-accretion.py:696-701 for an isothermal
    gas (c_eff = cs): min(dt_max, cfl dx / (max(cs + |v|) + 1e-30))."""
    cmax = np.max(cs + np.sqrt(vx * vx + vy * vy + vz * vz))
    return min(dt_max, cfl * dx / (cmax + 1e-30))


def apply_grav_accel(rho, vx, vy, vz, dt, g: Grid, ph: Physics):
    """quantum-fluid.py:267-278."""
    ax, ay, az = acceleration_grid(rho, g, ph)
    return vx + ax * dt, vy + ay * dt, vz + az * dt


def compute_step_gas(psi, rho, vx, vy, vz, t, dt, g: Grid, ph: Physics):
    """quantum-fluid.py:414-438 (cubic cells: dx for every axis)."""
    dx, vol = g.dx, g.vol
    rho_tot = np.abs(psi) ** 2 + rho
    V = get_potential(rho_tot, g, ph)
    psi = kick(psi, V, ph.m_per_hbar * dt / 2.0)
    vx, vy, vz = apply_grav_accel(rho_tot, vx, vy, vz, dt / 2.0, g, ph)
    psi = drift(psi, dt, g, ph)
    rho, vx, vy, vz = solve_hydro(rho, vx, vy, vz, dt, dx, vol, ph.cs, ph.flux)
    rho_tot = np.abs(psi) ** 2 + rho
    V = get_potential(rho_tot, g, ph)
    psi = kick(psi, V, ph.m_per_hbar * dt / 2.0)
    vx, vy, vz = apply_grav_accel(rho_tot, vx, vy, vz, dt / 2.0, g, ph)
    return psi, rho, vx, vy, vz, t + dt


# --------------------------------------------------------------------------
# Diagnostics (not in the reference: mass and energy for the error figures)
# --------------------------------------------------------------------------

def mass_total(psi, g: Grid, m_p=None, rho_gas=None):
    M = g.vol * np.sum(np.abs(psi) ** 2)
    if m_p is not None:
        M += float(np.sum(m_p))
    if rho_gas is not None:
        M += g.vol * np.sum(rho_gas)
    return M


def energy_total(psi, g: Grid, ph: Physics, pos=None, vel=None, m_p=None):
    """E = vol*sum(|grad psi|^2)/(2 mph^2) + vol/2*sum(V rho) + sum(m v^2)/2.

    Conserved by i d(psi)/dt = [-(1/(2 mph)) lap + mph V] psi with
    lap V = 4 pi G (rho - rho_bar) and dv/dt = -grad V.
    """
    m = make_grid(g)
    psi_hat = fftn(psi)
    K = g.vol * np.sum(m.k_sq * np.abs(psi_hat) ** 2) / g.ncell \
        / (2.0 * ph.m_per_hbar ** 2)
    rho = np.abs(psi) ** 2
    if pos is not None and len(pos):
        rho = rho + bin_particles(pos, m_p, g)
    ph0 = Physics(ph.m_per_hbar, ph.four_pi_G, ph.rho_bar, 0.0, ph.cs)
    V = get_potential(rho, g, ph0)
    W = 0.5 * g.vol * np.sum(V * rho)
    E = K + W
    if vel is not None and len(vel):
        mp = np.broadcast_to(np.asarray(m_p, dtype=np.float64), (vel.shape[0],))
        E += 0.5 * np.sum(mp[:, None] * vel ** 2)
    return E


# --------------------------------------------------------------------------
# Seeded initial conditions (the workload definitions; host FP64)
# --------------------------------------------------------------------------

def filtered_gaussian_field(g: Grid, seed: int, sigma_k: float = 16.0 * np.pi):
    """Re, Im ~ N(0,1) per cell (Re drawn first), multiplied in k-space by
    exp(-|k|^2/(2 sigma_k^2)); returns the complex filtered field."""
    rng = np.random.default_rng(seed)
    re = rng.standard_normal(g.shape)
    im = rng.standard_normal(g.shape)
    m = make_grid(g)
    f = np.fft.fftn(re + 1j * im)
    f *= np.exp(-m.k_sq / (2.0 * sigma_k ** 2))
    return np.fft.ifftn(f)


def filtered_gaussian_psi(g: Grid, seed: int, mean_rho: float = 1.0,
                          sigma_k: float = 16.0 * np.pi):
    """psi with mean |psi|^2 = mean_rho."""
    f = filtered_gaussian_field(g, seed, sigma_k)
    return f * np.sqrt(mean_rho / np.mean(np.abs(f) ** 2))


def star_ic(g: Grid, n_p: int, seed: int, total_mass: float = 0.1,
            sigma_v: float = 0.05):
    """Uniform FP64 positions in the box, N(0, sigma_v^2) velocities."""
    rng = np.random.default_rng(seed)
    L = np.array([g.Lx, g.Ly, g.Lz])
    pos = rng.random((n_p, 3)) * L
    vel = rng.normal(0.0, sigma_v, (n_p, 3))
    m_s = total_mass / n_p if n_p else 0.0
    return pos, vel, m_s


def gas_ic(g: Grid, seed: int, base: float = 0.2, amp: float = 0.1,
           floor: float = 0.01, sigma_k: float = 16.0 * np.pi):
    """rho_g = base*(1 + amp*g), g = Re of the filtered field scaled to zero
    mean and unit variance, clipped to rho_g >= floor; v = 0."""
    f = np.real(filtered_gaussian_field(g, seed, sigma_k))
    f = (f - f.mean()) / f.std()
    rho = np.maximum(base * (1.0 + amp * f), floor)
    z = np.zeros(g.shape)
    return rho, z.copy(), z.copy(), z.copy()


def default_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


# ---------------------------------------------------------------------------
# kick/drift variants: comoving cosmology and a static external potential
# ---------------------------------------------------------------------------
def compute_dt_cosmo(a, da, H0, omega_m, omega_l, n_quad=1000):
    """quantum-cosmo.py:149-163: dt between a and a+da from a midpoint
    quadrature of 1/adot over [0, a] and [0, a+da]."""
    h = 1.0 / n_quad
    lin = np.linspace(0.5 * h, 1.0 - 0.5 * h, n_quad)

    def t_of(aa):
        at = aa * lin
        adot = H0 * np.sqrt(omega_m / at + omega_l * at ** 2)
        return np.mean(1.0 / adot) * aa

    return t_of(a + da) - t_of(a)


def compute_step_cosmo(psi, a, da, g: Grid, ph: Physics, H0, omega_m, omega_l):
    """quantum-cosmo.py:166-197 (unmerged kicks): returns (psi, a + da)."""
    kin_fac = 0.5 * (a ** -2 + (a + da) ** -2)
    pot_fac = 0.5 * (a ** -1 + (a + da) ** -1)
    dt = compute_dt_cosmo(a, da, H0, omega_m, omega_l)
    V = get_potential(np.abs(psi) ** 2, g, ph)
    psi = kick(psi, V, ph.m_per_hbar * dt * pot_fac / 2.0)
    psi = drift(psi, dt * kin_fac, g, ph)
    V = get_potential(np.abs(psi) ** 2, g, ph)
    psi = kick(psi, V, ph.m_per_hbar * dt * pot_fac / 2.0)
    return psi, a + da


def external_point_mass(g: Grid, M, G):
    """quantum-tidal.py:175-179: -G M / (r + dx/2), r from the box centre."""
    X, Y, Z = cell_centres(g)
    dx = g.Lx / g.nx
    r = np.sqrt((X - 0.5 * g.Lx) ** 2 + (Y - 0.5 * g.Lx) ** 2 + (Z - 0.5 * g.Lx) ** 2)
    return -G * M / (r + 0.5 * dx)


def compute_step_tidal(psi, dt, V_ext, g: Grid, ph: Physics):
    """quantum-tidal.py:185-207 (unmerged kicks with V + V_ext)."""
    V = get_potential(np.abs(psi) ** 2, g, ph) + V_ext
    psi = kick(psi, V, ph.m_per_hbar * dt / 2.0)
    psi = drift(psi, dt, g, ph)
    V = get_potential(np.abs(psi) ** 2, g, ph) + V_ext
    return kick(psi, V, ph.m_per_hbar * dt / 2.0)


# ---------------------------------------------------------------------------
# black-hole accretion (quantum-bh.py)
# ---------------------------------------------------------------------------
def do_accretion(psi, pos, vel, m_bh, g: Grid, ph: Physics, G, c, sigma, dt):
    """quantum-bh.py:306-339 (returns new psi, m_bh)."""
    dx = g.Lx / g.nx
    dxs = np.array([dx, dx, dx])
    i = np.floor((pos - 0.5 * dxs) / dxs)
    i = np.mod(i, np.array([g.nx, g.ny, g.nz])).astype(int)
    s = 0
    psi_at_bh = psi[i[s, 0], i[s, 1], i[s, 2]]
    psi_amp = np.abs(psi_at_bh)
    psi_theta = np.angle(psi_at_bh)
    v = np.sqrt(sigma ** 2 + vel[0, 0] ** 2 + vel[0, 1] ** 2 + vel[0, 2] ** 2)
    xi = 2.0 * np.pi * G * m_bh[0] * ph.m_per_hbar / v
    dM_dt = (32.0 * np.pi ** 2 * (G * m_bh[0]) ** 3 * ph.m_per_hbar * psi_amp ** 2
             / (c ** 3 * v * (1.0 - np.exp(-xi))))
    new_amp = np.abs(psi_amp) - np.sqrt(dM_dt * dt / dx ** 3)
    psi = psi.copy()
    psi[i[s, 0], i[s, 1], i[s, 2]] = new_amp * np.exp(1.0j * psi_theta)
    return psi, m_bh + dM_dt * dt


def compute_step_bh(psi, m_bh, pos, vel, t, dt, g: Grid, ph: Physics, G, c, sigma):
    """quantum-bh.py:345-381: the KDK step with per-hole masses, then accretion."""
    psi, pos, vel, t, _, _ = compute_step_stars(psi, pos, vel, t, 0.0, np.inf, dt, m_bh, g, ph)
    psi, m_bh = do_accretion(psi, pos, vel, m_bh, g, ph, G, c, sigma, dt)
    return psi, m_bh, pos, vel, t


# ---------------------------------------------------------------------------
# black-hole force and ambient Bondi sink (quantum-fluid-accretion.py)
# ---------------------------------------------------------------------------
def lp_factor(g: Grid, k_cut):
    """quantum-fluid-accretion.py:143-145: exp(-(k^2 / (k_cut^2 + 1e-30))^4)."""
    m = make_grid(g)
    return np.exp(-(m.k_sq / (k_cut ** 2 + 1e-30)) ** 4)


def acceleration_grid_lp(rho, g: Grid, ph: Physics, k_cut):
    """accel_from_Vhat(V_hat * LP) with V_hat of get_potential_bg(rho,
    rho_bar) (quantum-fluid-accretion.py:182-191,719-722); (3, nx, ny, nz)."""
    m = make_grid(g)
    V_hat = _v_hat(rho, g, ph) * lp_factor(g, k_cut)
    ax = -np.real(ifftn(1.0j * m.kx * V_hat))
    ay = -np.real(ifftn(1.0j * m.ky * V_hat))
    az = -np.real(ifftn(1.0j * m.kz * V_hat))
    return np.stack((ax, ay, az))


def sample_cic(field, x, g: Grid):
    """_sample_cic of quantum-fluid-accretion.py:419-438 (cubic cells of
    size dx; field (nx, ny, nz), x one position)."""
    dx = g.Lx / g.nx
    n = g.nx
    xi = np.asarray(x, dtype=np.float64) / dx - 0.5
    i0 = np.floor(xi).astype(int)
    f = xi - i0
    ii = np.array([i0[0], i0[0] + 1, i0[0], i0[0], i0[0] + 1, i0[0] + 1, i0[0], i0[0] + 1])
    jj = np.array([i0[1], i0[1], i0[1] + 1, i0[1], i0[1] + 1, i0[1], i0[1] + 1, i0[1] + 1])
    kk = np.array([i0[2], i0[2], i0[2], i0[2] + 1, i0[2], i0[2] + 1, i0[2] + 1, i0[2] + 1])
    w = np.array([
        (1 - f[0]) * (1 - f[1]) * (1 - f[2]),
        f[0] * (1 - f[1]) * (1 - f[2]),
        (1 - f[0]) * f[1] * (1 - f[2]),
        (1 - f[0]) * (1 - f[1]) * f[2],
        f[0] * f[1] * (1 - f[2]),
        f[0] * (1 - f[1]) * f[2],
        (1 - f[0]) * f[1] * f[2],
        f[0] * f[1] * f[2],
    ])
    vals = field[np.mod(ii, n), np.mod(jj, n), np.mod(kk, n)]
    return np.sum(w * vals)


def bh_bondi_step_ambient(rho, M_bh_eff, x_bh, g: Grid, cs_iso, G, dt, r_acc_mult=2.5,
                          lam=float(np.exp(1.5) / 4.0), fmax=None, ann_lo=1.1, ann_hi=1.8,
                          vx=None, vy=None, vz=None, vxbh=0.0, vybh=0.0, vzbh=0.0):
    """quantum-fluid-accretion.py:440-510 (cubic box; defaults of :62-66).
    Returns (rho_new, dMdt_target, rho_inf, rB, dM_cap, r_acc, M_avail_kernel,
    dM_uncl)."""
    nx = rho.shape[0]
    dx = g.Lx / g.nx
    Lx = g.Lx
    xs = (np.arange(nx) + 0.5) * dx

    def pdelta(x0):  # :384-386
        d = np.abs(xs - x0)
        return np.minimum(d, Lx - d)

    dX, dY, dZ = np.meshgrid(pdelta(x_bh[0]), pdelta(x_bh[1]), pdelta(x_bh[2]), indexing="ij")
    R = np.sqrt(dX * dX + dY * dY + dZ * dZ)
    rB = G * M_bh_eff / (cs_iso * cs_iso + 1e-30)
    r_acc = r_acc_mult * dx
    rK = min(0.5 * r_acc, max(0.5 * dx, rB))
    mask_sink = R <= r_acc
    W_sink = np.exp(-0.5 * (R / (rK + 1e-30)) ** 2) * mask_sink
    r_lo = max(ann_lo * r_acc, r_acc + 0.25 * dx)
    r_hi = max(ann_hi * r_acc, r_lo + 0.50 * dx)
    shell = (R >= r_lo) & (R <= r_hi)
    rho_inf = np.sum(rho * shell) / (np.sum(shell) + 1e-30)
    if vx is not None and vy is not None and vz is not None:
        vx_inf = np.sum(vx * shell) / (np.sum(shell) + 1e-30)
        vy_inf = np.sum(vy * shell) / (np.sum(shell) + 1e-30)
        vz_inf = np.sum(vz * shell) / (np.sum(shell) + 1e-30)
        v_rel2 = (vxbh - vx_inf) ** 2 + (vybh - vy_inf) ** 2 + (vzbh - vz_inf) ** 2
        denom = (cs_iso ** 2 + v_rel2) ** 1.5
    else:
        denom = cs_iso ** 3
    dMdt_target = 4.0 * np.pi * lam * (G * M_bh_eff) ** 2 * rho_inf / (denom + 1e-30)
    dM_uncl = dMdt_target * dt
    WRho = W_sink * rho
    sumWRho = np.sum(WRho) + 1e-30
    phi = WRho / sumWRho
    M_avail_kernel = np.sum(rho * mask_sink) * dx ** 3
    dM_cap = min(dM_uncl, M_avail_kernel) if fmax is None else min(dM_uncl, fmax * M_avail_kernel)
    rho_new = np.maximum(0.0, rho - (dM_cap / dx ** 3) * phi)
    return rho_new, dMdt_target, rho_inf, rB, dM_cap, r_acc, M_avail_kernel, dM_uncl


def seed_mass_from_time(t, placed, init_m, inject_t, ramp_center=None, ramp_w=None):
    """quantum-fluid-accretion.py:682-690: init_m s(t) for a placed hole, with
    s = (1 + tanh((t - ramp_center) / (ramp_w + 1e-30))) / 2 for t >= inject_t
    (the ramp; ramp_center None = the step form s = [t >= inject_t])."""
    on = 1.0 if placed > 0.5 else 0.0
    if ramp_center is not None:
        s = 0.5 * (1.0 + np.tanh((t - ramp_center) / (ramp_w + 1e-30)))
        s = s * (1.0 if t >= inject_t else 0.0)
        return init_m * s * on
    return init_m * (1.0 if t >= inject_t else 0.0) * on
