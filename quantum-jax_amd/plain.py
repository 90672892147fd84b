"""Plain PyTorch restatement of the reference step (the "plain" GPU path).

This is the comparator of the benchmark metric ("c64 vs plain c128, same
GPU"): the reference operator sequence of quantum-stars.py:189-330 and
quantum-fluid.py:171-438 written with unfused torch ops (torch.fft on the
device, index_add_ scatter, gathers), in the reference's order with unmerged
half kicks.  It is selected explicitly; nothing in the HIP path falls back to
it.
"""

from __future__ import annotations

import math

import torch


class PlainKDK:
    def __init__(self, shape, L=(1.0, 1.0, 1.0), dtype=torch.complex128, device="cuda",
                 m_per_hbar=1.0, four_pi_G=1.0, rho_bar=1.0, cs=0.1):
        self.shape = tuple(shape)
        self.L = tuple(float(v) for v in L)
        self.cdt = dtype
        self.rdt = torch.float64 if dtype == torch.complex128 else torch.float32
        self.dev = torch.device(device)
        self.mph, self.G4, self.rho_bar, self.cs = m_per_hbar, four_pi_G, rho_bar, cs
        nx, ny, nz = self.shape
        self.d = tuple(Li / ni for Li, ni in zip(self.L, self.shape))
        self.vol = self.d[0] * self.d[1] * self.d[2]

        def kv(n, Ln):  # quantum-stars.py:142-148
            k = 2.0 * math.pi / Ln * torch.arange(-n / 2, n / 2, dtype=torch.float64)
            return torch.fft.ifftshift(k).to(self.rdt).to(self.dev)

        self.kx = kv(nx, self.L[0])[:, None, None]
        self.ky = kv(ny, self.L[1])[None, :, None]
        self.kz = kv(nz, self.L[2])[None, None, :]
        self.k_sq = self.kx ** 2 + self.ky ** 2 + self.kz ** 2
        self.k_sq_safe = self.k_sq + (self.k_sq == 0).to(self.rdt)

    # quantum-stars.py:189-193
    def get_potential(self, rho):
        V_hat = -torch.fft.fftn(self.G4 * (rho - self.rho_bar)) / self.k_sq_safe
        return torch.fft.ifftn(V_hat).real

    # quantum-stars.py:200-209
    def cic(self, pos):
        dxs = torch.tensor(self.d, dtype=torch.float64, device=self.dev)
        n = torch.tensor(self.shape, dtype=torch.float64, device=self.dev)
        i = torch.floor((pos - 0.5 * dxs) / dxs)
        ip1 = i + 1.0
        w_i = ((ip1 + 0.5) * dxs - pos) / dxs
        w_ip1 = (pos - (i + 0.5) * dxs) / dxs
        i = torch.remainder(i, n).long()
        ip1 = torch.remainder(ip1, n).long()
        return i, ip1, w_i, w_ip1

    _CORNERS = ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
                (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1))

    def _corners(self, pos):
        i, ip1, w_i, w_ip1 = self.cic(pos)
        nx, ny, nz = self.shape
        for sx, sy, sz in self._CORNERS:
            ix = (ip1 if sx else i)[:, 0]
            iy = (ip1 if sy else i)[:, 1]
            iz = (ip1 if sz else i)[:, 2]
            w = (w_ip1 if sx else w_i)[:, 0] * (w_ip1 if sy else w_i)[:, 1] \
                * (w_ip1 if sz else w_i)[:, 2]
            yield (ix * ny + iy) * nz + iz, w

    # quantum-stars.py:212-247 / quantum-bh.py:223-258
    def bin_particles(self, pos, m):
        rho = torch.zeros(math.prod(self.shape), dtype=self.rdt, device=self.dev)
        fac = m / self.vol
        for idx, w in self._corners(pos):
            rho.index_add_(0, idx, (w * fac).to(self.rdt))
        return rho.view(self.shape)

    # quantum-stars.py:250-288
    def get_acceleration(self, pos, rho):
        V_hat = -torch.fft.fftn(self.G4 * (rho - self.rho_bar)) / self.k_sq_safe
        ax = -torch.fft.ifftn(1j * self.kx * V_hat).real
        ay = -torch.fft.ifftn(1j * self.ky * V_hat).real
        az = -torch.fft.ifftn(1j * self.kz * V_hat).real
        a_grid = torch.stack((ax, ay, az), dim=-1)
        a_max = a_grid.abs().max()
        flat = a_grid.reshape(-1, 3)
        acc = torch.zeros(pos.shape[0], 3, dtype=torch.float64, device=self.dev)
        for idx, w in self._corners(pos):
            acc += w[:, None] * flat[idx].to(torch.float64)
        return acc, a_max

    def drift(self, psi, dt):
        psi_hat = torch.fft.fftn(psi)
        psi_hat = torch.exp(dt * (-1j * self.k_sq / self.mph / 2.0)) * psi_hat
        return torch.fft.ifftn(psi_hat)

    def kick(self, psi, V, dt):
        return torch.exp(-1j * self.mph * dt / 2.0 * V) * psi

    # quantum-stars.py:295-330
    def step_stars(self, psi, pos, vel, m_s, dt, a_max):
        L = torch.tensor(self.L, dtype=torch.float64, device=self.dev)
        rho_tot = psi.abs() ** 2 + self.bin_particles(pos, m_s)
        psi = self.kick(psi, self.get_potential(rho_tot), dt)
        acc, a1 = self.get_acceleration(pos, rho_tot)
        vel = vel + acc * dt / 2.0
        psi = self.drift(psi, dt)
        pos = torch.remainder(pos + vel * dt, L)
        rho_tot = psi.abs() ** 2 + self.bin_particles(pos, m_s)
        psi = self.kick(psi, self.get_potential(rho_tot), dt)
        acc, a2 = self.get_acceleration(pos, rho_tot)
        vel = vel + acc * dt / 2.0
        a_max = torch.maximum(torch.maximum(a_max, a1.to(torch.float64)), a2.to(torch.float64))
        dt_s = (self.d[0] / vel).abs().min()
        return psi, pos, vel, a_max, dt_s

    def step_plain(self, psi, dt):
        psi = self.kick(psi, self.get_potential(psi.abs() ** 2), dt)
        psi = self.drift(psi, dt)
        return self.kick(psi, self.get_potential(psi.abs() ** 2), dt)
