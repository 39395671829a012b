"""Pure-torch FP64 restatement of the RK4CN step, for reference gradients.

TEST INFRASTRUCTURE ONLY (imported by tests/ alone).  The same algorithm as
ns2d_oracle.rk4cn_step (reference RK4Step, Solver/src/solver.c:625-873, and
NonlinearRHSBatch, :884-1032) with the reference quirks off (no
ForceConjugacy write), written with differentiable torch operations so that
torch.autograd gives the exact gradient of the discrete map.  torch.fft's
irfft2/rfft2 have FFTW's conventions (unnormalised with norm="forward" /
"backward", imaginary parts of the ky=0 and ky=Ny/2 inputs ignored by irfft2).
"""
from __future__ import annotations

import math

import numpy as np
import torch

import ns2d_oracle as O


class TorchRK4CN:
    def __init__(self, p: O.Params):
        self.p = p
        Nx, Ny = p.Nx, p.Ny
        kx, ky = O.wavenumbers(Nx, Ny)
        k2 = kx * kx + ky * ky
        nz = k2 != 0
        inv = np.where(nz, 1.0 / np.where(nz, k2, 1), 0.0)
        t = lambda a: torch.tensor(np.broadcast_to(a, (Nx, Ny // 2 + 1)).copy())
        self.U = t(1j * ky * inv)           # u_hat = U w_hat
        self.V = t(-1j * kx * inv)          # v_hat = V w_hat
        self.X = t(np.where(nz, 1j * kx, 0.0) + 0j)
        self.Y = t(np.where(nz, 1j * ky, 0.0) + 0j)
        mask = O.dealias_factor(p) * np.where(nz, 1.0, 0.0) * (1.0 / (Nx * Ny)) ** 2
        self.M = t(mask + 0j)
        lam = O.visc_coeff(p, p.hyper_flag, "ab")
        self.lam = t(lam)

    def c2r(self, X):
        return torch.fft.irfft2(X, s=(self.p.Nx, self.p.Ny), norm="forward")

    def rhs(self, w):
        u = self.c2r(self.U * w)
        v = self.c2r(self.V * w)
        wx = self.c2r(self.X * w)
        wy = self.c2r(self.Y * w)
        return self.M * torch.fft.rfft2(u * wx + v * wy)

    def step(self, w, dt):
        K1 = self.rhs(w)
        K2 = self.rhs(w + (dt * 0.5) * K1)
        K3 = self.rhs(w + (dt * 0.5) * K2)
        K4 = self.rhs(w + dt * K3)
        D = dt * self.lam
        s = K1 / 6.0 + K2 / 3.0 + K3 / 3.0 + K4 / 6.0
        new = w * ((2.0 - D) / (2.0 + D)) + (2.0 * dt / (2.0 + D)) * s
        zero = torch.zeros_like(new)
        zero[0, 0] = 1.0
        return torch.where(zero.real > 0, torch.zeros_like(new), new)

    def enstrophy(self, w):
        """ComputeSystemMeasurables' enstrophy (sys_msr.c:192,220,360)."""
        Nx, Ny = self.p.Nx, self.p.Ny
        c = torch.full((1, Ny // 2 + 1), 2.0, dtype=torch.float64)
        c[0, 0] = 1.0
        c[0, -1] = 1.0
        a2 = w.real ** 2 + w.imag ** 2
        a2 = a2.clone()
        a2[0, 0] = 0.0
        return (c * a2).sum() * (4.0 * math.pi ** 2) * (0.5 / float(Nx * Ny) ** 2)
