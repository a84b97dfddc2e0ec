"""Seeded initial fields for the benchmark flows.

* ``random_solenoidal_field`` — divergence-free random-phase velocity with
  ``E(k) ∝ k⁴ exp(−2k²/k0²)`` and exact ``u_rms``
  (benchmarks/pending/dit_turbulence/run.py:95-125).  The complex Gaussian
  draws come from ``numpy.random.default_rng(seed)`` in the reference's order
  (real then imaginary part, component by component, on the half spectrum),
  so a seed gives the reference's field; the projection and inverse FFT run on
  the GPU in float64.
* ``taylor_green_2d`` — the D2Q9 Taylor-Green vortex
  (benchmarks/verified/taylor_green_2d/run.py:79-88).
* ``equilibrium_populations`` — populations at equilibrium with a given
  density and velocity, via the library's equilibrium kernel.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import ops

__all__ = ["random_solenoidal_field", "taylor_green_2d", "equilibrium_populations"]


def _wavenumbers(n: int, device) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    kx = torch.fft.rfftfreq(n, d=1.0 / n, dtype=torch.float64, device=device)
    k = torch.fft.fftfreq(n, d=1.0 / n, dtype=torch.float64, device=device)
    KZ, KY, KX = torch.meshgrid(k, k, kx, indexing="ij")
    return KZ, KY, KX


def random_solenoidal_field(n: int, k0: float = 4.0, u_rms: float = 0.03, seed: int = 0,
                            device: str | torch.device = "cuda:0", dtype: torch.dtype = torch.float32) -> torch.Tensor:
    """(3, n, n, n) velocity (x, y, z components; arrays indexed [z][y][x]), mean 0, ``u_rms`` exact."""
    n = int(n)
    rng = np.random.default_rng(seed)
    shape = (n, n, n // 2 + 1)
    KZ, KY, KX = _wavenumbers(n, device)
    K2 = KX * KX + KY * KY + KZ * KZ
    K = torch.sqrt(K2)
    spec = K**4 * torch.exp(-2.0 * (K / float(k0)) ** 2)
    amp = torch.where(K > 0, torch.sqrt(spec / (4.0 * math.pi * torch.where(K2 > 0, K2, torch.ones_like(K2)))),
                      torch.zeros_like(K))
    uh = torch.empty((3,) + shape, dtype=torch.complex128, device=device)
    for i in range(3):
        re = torch.from_numpy(rng.standard_normal(shape)).to(device)
        im = torch.from_numpy(rng.standard_normal(shape)).to(device)
        uh[i] = torch.complex(re, im) * amp
        del re, im
    kdotu = KX * uh[0] + KY * uh[1] + KZ * uh[2]
    fac = torch.where(K2 > 0, kdotu / torch.where(K2 > 0, K2, torch.ones_like(K2)), torch.zeros_like(kdotu))
    for i, kc in enumerate((KX, KY, KZ)):
        uh[i] -= kc * fac
    del kdotu, fac
    uh[:, K == 0] = 0.0
    u = torch.fft.irfftn(uh, s=(n, n, n), dim=(1, 2, 3))
    del uh
    u *= float(u_rms) / torch.sqrt(torch.mean(u * u))
    return u.to(dtype)


def taylor_green_2d(n: int, u0: float = 0.04, noise: float = 1e-4, seed: int = 0,
                    device: str | torch.device = "cuda:0", dtype: torch.dtype = torch.float32):
    """(ρ, u_x, u_y) on an n×n periodic grid: u_x = −u0 cos(kx) sin(ky), u_y = u0 sin(kx) cos(ky), k = 2π/n.

    ρ = 1 plus a seeded U(−noise, noise) perturbation (numpy ``default_rng(seed)``).
    """
    k = 2.0 * math.pi / n
    x = torch.arange(n, dtype=torch.float64, device=device)
    Y, X = torch.meshgrid(x, x, indexing="ij")
    ux = -u0 * torch.cos(k * X) * torch.sin(k * Y)
    uy = u0 * torch.sin(k * X) * torch.cos(k * Y)
    rng = np.random.default_rng(seed)
    rho = 1.0 + torch.from_numpy(rng.uniform(-noise, noise, size=(n, n))).to(device)
    return rho.to(dtype), ux.to(dtype), uy.to(dtype)


def equilibrium_populations(rho: torch.Tensor, u: torch.Tensor, lattice: str = "D3Q19") -> torch.Tensor:
    """Populations at equilibrium for density ρ (nz, ny, nx) and velocity u (3, nz, ny, nx)."""
    lat = lattice.upper()
    if lat == "D3Q19":
        return ops.equilibrium3d(rho, u[0], u[1], u[2])
    if lat == "D3Q27":
        return ops.equilibrium27(rho, u[0], u[1], u[2])
    if lat == "D2Q9":
        return ops.equilibrium(rho, u[0], u[1])
    raise ValueError(f"unknown lattice {lattice!r}")
