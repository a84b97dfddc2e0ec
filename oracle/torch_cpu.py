"""Eager-torch CPU restatement of the reference's D3Q19 channel step — TEST INFRASTRUCTURE ONLY.

The reference runs its per-step update on a CPU as a chain of eager torch ops:
``collide_bgk3d`` (macroscopic3d + equilibrium3d + BGK, solver3d.py:100-104 with
d3q19.py:83-186), ``stream3d_roll`` (19 ``torch.roll`` calls, solver3d.py:266-284),
``apply_simple_channel_boundaries_3d`` (equilibrium inlet with ρ copied from x=1,
zero-gradient outlet, full-grid macroscopic, boundaries3d.py:407-443) and the
obstacle's full-way swap ``bounce_back_cells_3d`` (boundaries3d.py:112-132).
This module restates that chain with the same tensor operations so that
``bench.py``'s ``cpu_baseline`` times the reference's CPU path, not a port of
the fused kernel.  ``tests/test_oracle_torch_cpu.py`` checks it against the
numpy oracle.

Only ``tests/`` and ``bench.py``'s ``cpu_baseline`` leg import it.
"""

from __future__ import annotations

import torch

from .lattice import D3Q19

_C = torch.tensor(D3Q19.c, dtype=torch.float32)  # (19, 3)
_W = torch.tensor(D3Q19.w, dtype=torch.float64)
_OPP = torch.tensor([int(v) for v in D3Q19.opposite], dtype=torch.long)


def macroscopic3d(f: torch.Tensor):
    """ρ = Σf, u = Σ c f / max(ρ, 1e-12) (d3q19.py:165-186)."""
    c = _C.to(f.dtype)
    rho = f.sum(dim=0)
    rho_safe = torch.clamp(rho, min=1e-12)
    ux = (f * c[:, 0].view(19, 1, 1, 1)).sum(dim=0) / rho_safe
    uy = (f * c[:, 1].view(19, 1, 1, 1)).sum(dim=0) / rho_safe
    uz = (f * c[:, 2].view(19, 1, 1, 1)).sum(dim=0) / rho_safe
    return rho, ux, uy, uz


def equilibrium3d(rho, ux, uy, uz):
    """w ρ (1 + 3 c·u + 4.5 (c·u)² − 1.5 u²) with 19-wide broadcasts (d3q19.py:83-125)."""
    c = _C.to(rho.dtype)
    w = _W.to(rho.dtype).view(19, 1, 1, 1)
    u_sq = ux * ux + uy * uy + uz * uz
    cu = c[:, 0].view(19, 1, 1, 1) * ux + c[:, 1].view(19, 1, 1, 1) * uy + c[:, 2].view(19, 1, 1, 1) * uz
    return w * rho.unsqueeze(0) * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u_sq.unsqueeze(0))


def collide_bgk3d(f: torch.Tensor, tau: float) -> torch.Tensor:
    """f − (f − f_eq)/τ (solver3d.py:100-104)."""
    feq = equilibrium3d(*macroscopic3d(f))
    return f - (f - feq) / tau


def stream3d_roll(f: torch.Tensor) -> torch.Tensor:
    """Periodic pull, one ``torch.roll`` per direction (solver3d.py:266-284)."""
    out = torch.empty_like(f)
    for q in range(19):
        cx, cy, cz = (int(v) for v in D3Q19.c[q])
        out[q] = torch.roll(f[q], shifts=(cz, cy, cx), dims=(0, 1, 2))
    return out


def bounce_back_cells_3d(f: torch.Tensor, mask: torch.Tensor) -> torch.Tensor:
    """where(mask, f[opp], f) (boundaries3d.py:112-132)."""
    return torch.where(mask.unsqueeze(0), f[_OPP], f)


def channel_boundaries(f: torch.Tensor, u_in: float, obstacle: torch.Tensor) -> torch.Tensor:
    """Equilibrium inlet (ρ from x=1), copy outlet, obstacle bounce-back (boundaries3d.py:407-443)."""
    rho, ux, uy, uz = macroscopic3d(f)
    ux[:, :, 0] = u_in
    uy[:, :, 0] = 0.0
    uz[:, :, 0] = 0.0
    rho[:, :, 0] = rho[:, :, 1]
    feq_in = equilibrium3d(rho[:, :, 0:1], ux[:, :, 0:1], uy[:, :, 0:1], uz[:, :, 0:1])
    f[:, :, :, 0] = feq_in[:, :, :, 0]
    f[:, :, :, -1] = f[:, :, :, -2]
    return bounce_back_cells_3d(f, obstacle)


def channel_step(f: torch.Tensor, tau: float, u_in: float, obstacle: torch.Tensor) -> torch.Tensor:
    """One eager step: collide, stream, channel boundaries and obstacle bounce-back."""
    return channel_boundaries(stream3d_roll(collide_bgk3d(f, tau)), u_in, obstacle)
