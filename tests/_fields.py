"""Seeded initial fields shared by the parity tests (numpy, float64)."""

from __future__ import annotations

import numpy as np

from oracle import ref
from oracle.lattice import D2Q9, D3Q19, D3Q27

LATS = {9: D2Q9, 19: D3Q19, 27: D3Q27}


def random_state(q: int, shape, seed: int = 0, noise: float = 1e-3, u_amp: float = 0.04) -> np.ndarray:
    """Equilibrium of a random rho/u field plus seeded non-equilibrium noise."""
    rng = np.random.default_rng(seed)
    lat = LATS[q]
    sh = tuple(shape)
    rho = 1.0 + 0.02 * rng.standard_normal(sh)
    u = u_amp * rng.standard_normal((3,) + sh)
    if q == 9:
        u[2] = 0.0
    f = ref.equilibrium(lat, rho, u[0], u[1], u[2])
    return f + noise * rng.standard_normal(f.shape) * lat.w.reshape((-1,) + (1,) * len(sh))


def shear_yz(n: int) -> np.ndarray:
    """ux(z), uy(z), uz(y) shear (the reference's z-asymmetric field)."""
    ax = np.arange(n) / n
    ux = np.broadcast_to((0.08 * np.sin(2 * np.pi * ax + 0.3))[:, None, None], (n, n, n))
    uy = np.broadcast_to((0.06 * np.sin(2 * np.pi * ax + 1.1))[:, None, None], (n, n, n))
    uz = np.broadcast_to((0.05 * np.sin(2 * np.pi * ax + 0.7))[None, :, None], (n, n, n))
    return ref.equilibrium(D3Q19, np.ones((n, n, n)), ux.copy(), uy.copy(), uz.copy())


def blob_offcentre(n: int) -> np.ndarray:
    a = np.arange(n, dtype=np.float64)
    z, y, x = np.meshgrid(a, a, a, indexing="ij")
    s = n / 8.0
    r2 = (z - 0.31 * n) ** 2 + (y - 0.43 * n) ** 2 + (x - 0.62 * n) ** 2
    rho = 1.0 + 0.4 * np.exp(-r2 / (2 * s * s))
    d = 0.02
    return ref.equilibrium(D3Q19, rho, np.full_like(rho, d), np.full_like(rho, -d), np.full_like(rho, d))


def random_eq(n: int, seed: int = 20260819) -> np.ndarray:
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.05 * rng.standard_normal((n, n, n))
    u = 0.03 * rng.standard_normal((3, n, n, n))
    return ref.equilibrium(D3Q19, rho, u[0], u[1], u[2])


ASYM_FIELDS = {"shear_yz": shear_yz, "blob_offcentre": blob_offcentre, "random": random_eq}


def canary(q: int, nz: int, ny: int, nx: int) -> np.ndarray:
    """Integer-encoded field q*2^18 + z*2^12 + y*2^6 + x (exact in float32)."""
    qi = np.arange(q).reshape(q, 1, 1, 1)
    zi = np.arange(nz).reshape(1, nz, 1, 1)
    yi = np.arange(ny).reshape(1, 1, ny, 1)
    xi = np.arange(nx).reshape(1, 1, 1, nx)
    return (qi * 2**18 + zi * 2**12 + yi * 2**6 + xi).astype(np.float64)
