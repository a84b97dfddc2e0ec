"""Per-field comparisons of population states (ρ, u, Π), computed on the GPU.

Used to quote the error of a reduced-precision storage run against an FP64 run of the
same flow: relative L2 errors of ρ − 1, of u, and of the deviatoric second moment
Π_ab − ρ c_s² δ_ab = Σ_q c_qa c_qb f_q − ρ/3 δ_ab.
"""

from __future__ import annotations

import torch

from . import lattice as _lat
from . import ops

__all__ = ["moments", "field_errors"]


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
