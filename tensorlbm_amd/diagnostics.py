"""Per-field comparisons of population states (ρ, u, Π), computed on the GPU, and drag normalisation.

Used to quote the error of a reduced-precision storage run against an FP64 run of the
same flow: relative L2 errors of ρ − 1, of u, and of the deviatoric second moment
Π_ab − ρ c_s² δ_ab = Σ_q c_qa c_qb f_q − ρ/3 δ_ab.

:func:`obstacle_drag_coefficient` turns the fused step's momentum-exchange force into
C_D = F_x / (½ ρ u_in² S_ref) with the wetted voxel area (or the frontal area, or a
given one) as S_ref (triton_fused_obstacle.py:2156-2262).  The areas are counted once
per geometry, on the device the mask lives on.
"""

from __future__ import annotations

import torch

from . import lattice as _lat
from . import ops

__all__ = ["moments", "field_errors", "voxel_wetted_area", "frontal_area", "obstacle_drag_coefficient"]


def moments(f: torch.Tensor, *, deviation: bool = False):
    """(ρ, u (3, ...), Π_dev (6, ...): xx yy zz xy xz yz) in float64 from (Q, nz, ny, nx) populations."""
    q = int(f.shape[0])
    tab = _lat.BY_Q[q]
    f = f.to(torch.float64)
    if deviation:
        f = f + tab.W_EXACT64.to(f.device).reshape((q,) + (1,) * (f.ndim - 1))
    f = f.contiguous()
    if q == 19:
        rho, ux, uy, uz = ops.macroscopic3d(f)
    elif q == 27:
        rho, ux, uy, uz = ops.macroscopic27(f)
    else:
        raise ValueError("3-D lattices only")
    c = tab.C.to(device=f.device, dtype=torch.float64)
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    pi = []
    flat = f.reshape(q, -1)
    for a, b in pairs:
        m = (c[:, a] * c[:, b]) @ flat
        if a == b:
            m = m - rho.reshape(-1) / 3.0
        pi.append(m.reshape(rho.shape))
    return rho, torch.stack([ux, uy, uz]), torch.stack(pi)


def _rel(a: torch.Tensor, b: torch.Tensor) -> float:
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b).clamp_min(1e-300))


def field_errors(f: torch.Tensor, f_ref: torch.Tensor, *, deviation: bool = False,
                 ref_deviation: bool = False) -> dict:
    """Relative L2 errors of ρ − 1, u and Π_dev of ``f`` against ``f_ref``."""
    r, u, p = moments(f, deviation=deviation)
    r0, u0, p0 = moments(f_ref, deviation=ref_deviation)
    return {"rho": _rel(r - 1.0, r0 - 1.0), "u": _rel(u, u0), "pi": _rel(p, p0),
            "u_max_abs": float((u - u0).abs().max())}


def _solid3d(mask: torch.Tensor) -> torch.Tensor:
    if not isinstance(mask, torch.Tensor) or mask.ndim != 3:
        raise ValueError("mask must be a 3D tensor (nz, ny, nx)")
    return mask if mask.dtype == torch.bool else mask != 0


def voxel_wetted_area(mask: torch.Tensor, dx: float = 1.0) -> float:
    """Area of the solid cells' faces that touch a fluid cell or the domain's outside, times dx².

    With the mask padded by one fluid layer, every such face is exactly one
    solid/fluid change between neighbours along one axis.
    """
    padded = torch.nn.functional.pad(_solid3d(mask).to(torch.uint8), (1, 1, 1, 1, 1, 1))
    faces = sum(int((padded.diff(dim=d) != 0).sum()) for d in range(3))
    return float(faces) * dx * dx


def frontal_area(mask: torch.Tensor, streamwise_axis: int = 2, dx: float = 1.0) -> float:
    """Projected area: the number of stream-wise cell columns holding any solid cell, times dx²."""
    if streamwise_axis not in (0, 1, 2):
        raise ValueError(f"streamwise_axis must be 0, 1 or 2, got {streamwise_axis}")
    return float(int(_solid3d(mask).any(dim=streamwise_axis).sum())) * dx * dx


def obstacle_drag_coefficient(fx: float, obstacle: torch.Tensor, u_in: float, *, dx: float = 1.0,
                              rho_lu: float = 1.0, s_ref: str = "wetted", streamwise_axis: int = 2,
                              s_ref_value: float | None = None) -> dict:
    """C_D = F_x / (½ ρ u_in² S_ref) from a stream-wise momentum-exchange force (triton_fused_obstacle.py:2200-2262).

    ``S_ref`` is ``s_ref_value`` when given, else the wetted voxel area
    (``s_ref="wetted"``, the reference's production convention) or the frontal
    area across ``streamwise_axis`` (``"frontal"``).  Returns ``C_D``, ``S_ref``,
    ``s_ref_kind``, ``dynamic_pressure`` and ``fx``; C_D is 0 when the dynamic
    pressure is 0.
    """
    if s_ref_value is not None:
        kind, area = "explicit", float(s_ref_value)
    elif s_ref == "wetted":
        kind, area = "wetted", voxel_wetted_area(obstacle, dx)
    elif s_ref == "frontal":
        kind, area = "frontal", frontal_area(obstacle, streamwise_axis, dx)
    else:
        raise ValueError(f"s_ref must be 'wetted' or 'frontal', got {s_ref!r}")
    q = 0.5 * rho_lu * u_in * u_in * area
    return {"C_D": float(fx) / q if q > 0.0 else 0.0, "S_ref": area, "s_ref_kind": kind, "dynamic_pressure": q,
            "fx": float(fx)}

