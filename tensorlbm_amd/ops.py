"""Reference-named functional ops, each one HIP launch on the caller's stream.

Signatures, argument meaning and error behaviour follow the reference:

* D2Q9  ``equilibrium``/``macroscopic`` (d2q9.py:39-74), ``collide_bgk``,
  ``collide_mrt``, ``stream``, ``correct_mass`` (solver.py:67-437),
  ``bounce_back_cells`` (boundaries.py:39-47).
* D3Q19 ``equilibrium3d``/``macroscopic3d`` (d3q19.py:83-186),
  ``collide_bgk3d``/``collide_mrt3d``/``stream3d``/``stream3d_roll``/
  ``correct_mass3d`` (solver3d.py:100-300), ``bounce_back_cells_3d``
  (boundaries3d.py:112-132), ``compute_obstacle_forces_3d`` (obstacles.py:91-130).
* D3Q27 ``equilibrium27``/``macroscopic27``/``collide_bgk27``/``collide_mrt27``/
  ``stream27``/``correct_mass27`` (d3q27.py:353-791).
* Collision families ``collide_cascaded_d3q19/_d3q27`` (cascaded_collision.py:445-),
  ``collide_cumulant_d3q19`` (cumulant.py:1159-1377),
  ``collide_cumulant_geier_d3q27`` (cumulant.py:1050-1156),
  ``collide_smagorinsky_bgk/_bgk3d/_bgk27`` (turbulence.py:216-494).

Every op returns a new tensor (or writes ``out=``) and never mutates its
inputs.  Inputs must be GPU tensors; float32 and float64 are supported for
the eager ops (the fused solver also takes 16-bit storage).
"""

from __future__ import annotations

import ctypes
import math

import torch

from . import _native as nat
from .lattice import BY_Q

__all__ = [
    "equilibrium", "macroscopic", "equilibrium3d", "equilibrium3d_low_memory", "macroscopic3d",
    "macroscopic3d_low_memory", "equilibrium27", "macroscopic27", "collide_bgk", "collide_bgk3d",
    "collide_bgk27", "collide_mrt", "collide_mrt3d", "collide_mrt27", "collide_cascaded_d3q19",
    "collide_cascaded_d3q27", "collide_cumulant_d2q9", "collide_cumulant_d3q19", "collide_cumulant_d3q27",
    "collide_cumulant_geier_d3q27",
    "collide_smagorinsky_bgk", "collide_smagorinsky_bgk3d", "collide_smagorinsky_bgk27", "collide_smagorinsky_mrt",
    "collide_smagorinsky_mrt3d", "collide_smagorinsky_mrt27", "stream",
    "stream3d", "stream3d_roll", "stream27", "stream27_roll", "correct_mass", "correct_mass3d",
    "correct_mass27", "bounce_back_cells", "bounce_back_cells_3d", "bounce_back_cells_27",
    "compute_obstacle_forces_3d", "compute_obstacle_forces_27", "moment_sums", "gradient_closure_tau", "collide_bgk_fused", "collide_bgk_matmul",
    "apply_simple_channel_boundaries_3d", "far_field_bc_3d", "zou_he_moving_lid_3d", "sphere_mask",
]

_EAGER = (torch.float32, torch.float64)


def _eager_dtype(t: torch.Tensor, what: str) -> int:
    if t.dtype not in _EAGER:
        raise ValueError(f"Unsupported dtype {t.dtype} for {what}; use float32 or float64")
    return nat.STORE_F32 if t.dtype == torch.float32 else nat.STORE_F64


def _grid(f: torch.Tensor, q: int, what: str) -> tuple[int, int, int]:
    if f.ndim == 3 and q == 9:
        return 1, int(f.shape[1]), int(f.shape[2])
    if f.ndim == 4 and q in (19, 27):
        return int(f.shape[1]), int(f.shape[2]), int(f.shape[3])
    want = "(9, ny, nx)" if q == 9 else f"({q}, nz, ny, nx)"
    raise ValueError(f"{what}: populations must have shape {want}, got {tuple(f.shape)}")


def _pops(f: torch.Tensor, q: int, what: str, contiguous: bool = True) -> tuple[torch.Tensor, tuple[int, int, int]]:
    nat.require_gpu(f, what)
    if f.shape[0] != q:
        raise ValueError(f"{what}: expected Q={q} populations, got shape {tuple(f.shape)}")
    dims = _grid(f, q, what)
    if contiguous or f.stride(-1) != 1:
        f = f.contiguous()
    return f, dims


# ---------------------------------------------------------------------------
# equilibrium / macroscopic
# ---------------------------------------------------------------------------


def _equilibrium(q, rho, ux, uy, uz, out, what):
    fields = (rho, ux, uy) if uz is None else (rho, ux, uy, uz)
    if not all(t.shape == rho.shape for t in fields):
        names = "rho, ux, and uy" if uz is None else "rho, ux, uy, and uz"
        raise ValueError(f"{names} shapes must match: " + ", ".join(str(tuple(t.shape)) for t in fields))
    nat.require_gpu(rho, what)
    dt = _eager_dtype(rho, what)
    fields = [t.to(device=rho.device, dtype=rho.dtype).contiguous() for t in fields]
    shape = (q,) + tuple(rho.shape)
    res = out if out is not None else torch.empty(shape, dtype=rho.dtype, device=rho.device)
    if tuple(res.shape) != shape or res.dtype != rho.dtype or not res.is_contiguous():
        raise ValueError(f"out must be a contiguous {rho.dtype} tensor of shape {shape}")
    uzp = nat.ptr(fields[3]) if uz is not None else None
    rc = nat.lib().tlbm_equilibrium(q, dt, nat.ptr(fields[0]), nat.ptr(fields[1]), nat.ptr(fields[2]), uzp,
                                    res.data_ptr(), rho.numel(), nat.stream_of(rho))
    nat.check(rc, what)
    return res


def equilibrium(rho, ux, uy, device=None):
    """D2Q9 f_eq for rho and velocity fields (d2q9.py:39-58)."""
    del device
    return _equilibrium(9, rho, ux, uy, None, None, "equilibrium")


def equilibrium3d(rho, ux, uy, uz, device=None, *, out=None):
    """D3Q19 f_eq of shape (19, nz, ny, nx) (d3q19.py:83-125); ``out=`` is filled in place."""
    del device
    return _equilibrium(19, rho, ux, uy, uz, out, "equilibrium3d")


def equilibrium3d_low_memory(rho, ux, uy, uz, device=None, *, out=None):
    """Same values as :func:`equilibrium3d` (d3q19.py:128-162); the kernel never materialises temporaries."""
    return equilibrium3d(rho, ux, uy, uz, device, out=out)


def equilibrium27(rho, ux, uy, uz, device=None, *, out=None):
    """D3Q27 f_eq (d3q27.py:353-392)."""
    del device
    return _equilibrium(27, rho, ux, uy, uz, out, "equilibrium27")


def _macroscopic(f, q, what):
    f, dims = _pops(f, q, what)
    dt = _eager_dtype(f, what)
    spatial = tuple(f.shape[1:])
    rho, ux, uy = (torch.empty(spatial, dtype=f.dtype, device=f.device) for _ in range(3))
    uz = torch.empty(spatial, dtype=f.dtype, device=f.device) if q != 9 else None
    rc = nat.lib().tlbm_macroscopic(q, dt, f.data_ptr(), rho.data_ptr(), ux.data_ptr(), uy.data_ptr(),
                                    nat.ptr(uz), dims[0] * dims[1] * dims[2], nat.stream_of(f))
    nat.check(rc, what)
    return (rho, ux, uy) if q == 9 else (rho, ux, uy, uz)


def macroscopic(f, device=None):
    """D2Q9 (rho, ux, uy) (d2q9.py:61-74)."""
    del device
    return _macroscopic(f, 9, "macroscopic")


def macroscopic3d(f, device=None):
    """D3Q19 (rho, ux, uy, uz) (d3q19.py:165-186)."""
    del device
    return _macroscopic(f, 19, "macroscopic3d")


def macroscopic3d_low_memory(f):
    """D3Q19 moments (d3q19.py:189-212); same kernel as :func:`macroscopic3d`."""
    if not isinstance(f, torch.Tensor) or f.ndim != 4 or f.shape[0] != 19:
        raise ValueError("f must have shape (19,nz,ny,nx)")
    return _macroscopic(f, 19, "macroscopic3d_low_memory")


def macroscopic27(f, device=None):
    """D3Q27 (rho, ux, uy, uz) (d3q27.py:395-420)."""
    del device
    return _macroscopic(f, 27, "macroscopic27")


# ---------------------------------------------------------------------------
# collisions (local, one launch)
# ---------------------------------------------------------------------------


def _collide(f, q, coll, what):
    f, dims = _pops(f, q, what)
    dt = _eager_dtype(f, what)
    out = torch.empty_like(f)
    rc = nat.lib().tlbm_collide(q, dt, f.data_ptr(), None, out.data_ptr(), None, *dims, coll, nat.stream_of(f))
    nat.check(rc, what)
    return out


def _tau(tau, what):
    if isinstance(tau, torch.Tensor):
        raise NotImplementedError(f"{what}: tensor-valued tau (differentiable path) is not on the HIP path")
    return float(tau)


def collide_bgk(f, tau):
    """D2Q9 BGK ``f − (f − f_eq)/τ`` (solver.py:67-71)."""
    return _collide(f, 9, nat.make_collision(nat.COLL_BGK, _tau(tau, "collide_bgk"), eq_rho_floor=False),
                    "collide_bgk")


def collide_bgk_fused(f, tau):
    """D2Q9 BGK in one pass (solver.py:74-117): the same update as ``collide_bgk``, one kernel here too."""
    return _collide(f, 9, nat.make_collision(nat.COLL_BGK, _tau(tau, "collide_bgk_fused"), eq_rho_floor=False),
                    "collide_bgk_fused")


def collide_bgk_matmul(f, tau):
    """D2Q9 BGK with moments by matrix products (solver.py:194-246): the same update as ``collide_bgk``;
    the moments of nine populations are a few adds per cell, no GEMM."""
    return _collide(f, 9, nat.make_collision(nat.COLL_BGK, _tau(tau, "collide_bgk_matmul"), eq_rho_floor=False),
                    "collide_bgk_matmul")


def collide_bgk3d(f, tau):
    """D3Q19 BGK (solver3d.py:100-104)."""
    return _collide(f, 19, nat.make_collision(nat.COLL_BGK, _tau(tau, "collide_bgk3d"), eq_rho_floor=False),
                    "collide_bgk3d")


def collide_bgk27(f, tau):
    """D3Q27 BGK (d3q27.py:423-435)."""
    return _collide(f, 27, nat.make_collision(nat.COLL_BGK, _tau(tau, "collide_bgk27"), eq_rho_floor=False),
                    "collide_bgk27")


def collide_mrt(f, tau, s_e=1.64, s_eps=1.54, s_q=1.7, tau_field=None):
    """D2Q9 MRT (solver.py:249-335); ``tau_field`` gives a per-cell shear τ."""
    tau = _tau(tau, "collide_mrt")
    mode, te = nat.TAU_CONST, None
    if tau_field is not None:
        if tuple(tau_field.shape) != tuple(f.shape[1:]):
            raise ValueError(f"tau_field shape {tuple(tau_field.shape)} must match {tuple(f.shape[1:])}")
        mode, te = nat.TAU_FIELD, tau_field.to(device=f.device, dtype=torch.float32).contiguous()
    coll = nat.make_collision(nat.COLL_MRT, tau, tau_mode=mode, eq_rho_floor=False,
                              rates=(s_e, s_eps, s_q, 0.0), tau_eff=te)
    return _collide(f, 9, coll, "collide_mrt")


def collide_mrt3d(f, tau, s_e=1.19, s_eps=1.4, s_q=1.2, s_pi=None):
    """D3Q19 MRT (solver3d.py:107-160); ``s_pi`` defaults to ``s_e``."""
    s_pi = s_e if s_pi is None else s_pi
    coll = nat.make_collision(nat.COLL_MRT, _tau(tau, "collide_mrt3d"), eq_rho_floor=False,
                              rates=(s_e, s_eps, s_q, s_pi))
    return _collide(f, 19, coll, "collide_mrt3d")


def collide_mrt27(f, tau, s_e=1.19, s_eps=1.4, s_q=1.2, s_pi=None):
    """D3Q27 MRT, Gram-Schmidt basis (d3q27.py:643-726)."""
    s_pi = s_e if s_pi is None else s_pi
    coll = nat.make_collision(nat.COLL_MRT, _tau(tau, "collide_mrt27"), eq_rho_floor=False,
                              rates=(s_e, s_eps, s_q, s_pi))
    return _collide(f, 27, coll, "collide_mrt27")


def collide_cumulant_d2q9(f, tau, omega_b=1.0, omega_3=1.0, omega_4=1.0):
    """D2Q9 central-moment ("cumulant") collision (cumulant.py:116-311).

    Deviatoric stress at ω = 1/τ, trace at ``omega_b``, κ21/κ12 at ``omega_3``, κ22 at ``omega_4``;
    equilibrium central moments from the discrete equilibrium, so f_eq is a fixed point.
    """
    if f.ndim != 3 or f.shape[0] != 9:
        raise ValueError(f"D2Q9 populations must have shape (9, ny, nx), got {tuple(f.shape)}")
    tau = _tau(tau, "collide_cumulant_d2q9")
    coll = nat.make_collision(nat.COLL_CM, tau, eq_rho_floor=False, rates=(omega_b, omega_3, omega_4))
    return _collide(f, 9, coll, "collide_cumulant_d2q9")


def collide_cascaded_d3q19(f, tau, s_bulk=1.0, s_3=1.0, s_4=1.0):
    """Cascaded central-moment collision (cascaded_collision.py:445-509)."""
    if f.ndim != 4 or f.shape[0] != 19:
        raise ValueError(f"D3Q19 populations must have shape (19, nz, ny, nx), got {tuple(f.shape)}")
    tau = _tau(tau, "collide_cascaded_d3q19")
    if tau <= 0.5:
        raise ValueError(f"tau must be > 0.5, got {tau}")
    coll = nat.make_collision(nat.COLL_CM, tau, eq_rho_floor=False, rates=(s_bulk, s_3, s_4, s_4, s_4))
    return _collide(f, 19, coll, "collide_cascaded_d3q19")


def collide_cascaded_d3q27(f, tau, s_bulk=1.0, s_3=1.0, s_4=1.0, s_5=None, s_6=None):
    """D3Q27 cascaded collision; ``s_5``/``s_6`` default to ``s_4``."""
    if f.ndim != 4 or f.shape[0] != 27:
        raise ValueError(f"D3Q27 populations must have shape (27, nz, ny, nx), got {tuple(f.shape)}")
    tau = _tau(tau, "collide_cascaded_d3q27")
    if tau <= 0.5:
        raise ValueError(f"tau must be > 0.5, got {tau}")
    s_5 = s_4 if s_5 is None else s_5
    s_6 = s_4 if s_6 is None else s_6
    coll = nat.make_collision(nat.COLL_CM, tau, eq_rho_floor=False, rates=(s_bulk, s_3, s_4, s_5, s_6))
    return _collide(f, 27, coll, "collide_cascaded_d3q27")


def collide_cumulant_d3q19(f, tau, omega_b=1.0, omega_odd=1.0, omega_even=1.0, C_s=0.0, C_w=0.0, C_v=0.0,
                           solid_mask=None, wall_velocity=(0.0, 0.0, 0.0)):
    """D3Q19 Hermite-regularised cumulant (cumulant.py:1159-1377).

    ``C_s > 0`` applies that function's unclamped per-cell Smagorinsky τ_eff.
    ``C_w`` / ``C_v`` apply the WALE / Vreman closures: τ_eff from the velocity
    gradient (``gradient_closure_tau``), with ``wall_velocity`` on ``solid_mask``.
    """
    if min(C_s, C_w, C_v) < 0.0:
        raise ValueError("SGS coefficients must be non-negative")
    if sum(v > 0.0 for v in (C_s, C_w, C_v)) > 1:
        raise ValueError("only one SGS model may be active")
    if len(wall_velocity) != 3 or not all(math.isfinite(float(v)) for v in wall_velocity):
        raise ValueError("wall_velocity must contain three finite values")
    if solid_mask is not None and (tuple(solid_mask.shape) != tuple(f.shape[1:]) or solid_mask.dtype is not torch.bool
                                   or solid_mask.device != f.device):
        raise ValueError("solid_mask must be bool with the population spatial shape")
    tau = _tau(tau, "collide_cumulant_d3q19")
    rates = (omega_b, omega_odd, omega_even)
    if C_w > 0.0 or C_v > 0.0:
        te = gradient_closure_tau(f, tau, C_w=C_w, C_v=C_v, solid_mask=solid_mask, wall_velocity=wall_velocity)
        mode = nat.TAU_FIELD64 if te.dtype == torch.float64 else nat.TAU_FIELD
        coll = nat.make_collision(nat.COLL_CUMULANT, tau, tau_mode=mode, eq_rho_floor=False, rates=rates, tau_eff=te)
        out = _collide(f, 19, coll, "collide_cumulant_d3q19")
        del te
        return out
    mode = nat.TAU_SMAG_RAW if C_s > 0.0 else nat.TAU_CONST
    coll = nat.make_collision(nat.COLL_CUMULANT, tau, tau_mode=mode, eq_rho_floor=False, cs_delta_sq=C_s * C_s,
                              rates=rates)
    return _collide(f, 19, coll, "collide_cumulant_d3q19")


def gradient_closure_tau(f, tau, *, C_w=0.0, C_v=0.0, solid_mask=None, wall_velocity=(0.0, 0.0, 0.0)):
    """Per-cell τ_eff of the WALE (``C_w``) or Vreman (``C_v``) closure of a D3Q19 state.

    τ_eff = clip(τ + 3 ν_t, 0.5001, max(1, τ + 0.5)) (turbulence.py:875-888), ν_t from the
    macroscopic velocity with ``wall_velocity`` on ``solid_mask`` (cumulant.py:1263-1303),
    differentiated centrally inside and one-sidedly on the faces (turbulence.py:595-695).
    Returns a (nz, ny, nx) tensor in the populations' dtype (float32 or float64): the
    ``tau_eff`` input of ``collide_*`` and of the fused step (``TLBM_TAU_FIELD``).
    """
    if (C_w > 0.0) == (C_v > 0.0):
        raise ValueError("exactly one of C_w (WALE) and C_v (Vreman) must be positive")
    if min(C_w, C_v) < 0.0:
        raise ValueError("SGS coefficients must be non-negative")
    fc, dims = _pops(f, 19, "gradient_closure_tau")
    dt = _eager_dtype(fc, "gradient_closure_tau")
    spatial = tuple(fc.shape[1:])
    rho, ux, uy, uz = (torch.empty(spatial, dtype=fc.dtype, device=fc.device) for _ in range(4))
    s = nat.stream_of(fc)
    nat.check(nat.lib().tlbm_macroscopic(19, dt, fc.data_ptr(), rho.data_ptr(), ux.data_ptr(), uy.data_ptr(),
                                         uz.data_ptr(), fc[0].numel(), s), "gradient_closure_tau")
    mask = None
    if solid_mask is not None:
        if tuple(solid_mask.shape) != spatial:
            raise ValueError("solid_mask must have the population spatial shape")
        mask = solid_mask.to(device=fc.device, dtype=torch.int8).contiguous()
    wall = (ctypes.c_double * 3)(*(float(v) for v in wall_velocity))
    te = torch.empty(spatial, dtype=fc.dtype, device=fc.device)
    model, coef = (nat.SGS_WALE, C_w) if C_w > 0.0 else (nat.SGS_VREMAN, C_v)
    nat.check(nat.lib().tlbm_gradient_closure_tau(dt, ux.data_ptr(), uy.data_ptr(), uz.data_ptr(), nat.ptr(mask), wall,
                                                  *dims, model, float(coef), float(tau), te.data_ptr(), s),
              "gradient_closure_tau")
    del rho
    return te


def collide_cumulant_d3q27(f, tau, omega_b=1.0, omega_odd=1.0, omega_even=1.0, C_s=0.0):
    """Legacy D3Q27 central-moment operator with the pseudo-cumulant map on f_neq (cumulant.py:935-1047).

    ``C_s > 0`` relaxes every cell with one rate: the domain mean of the per-cell clamped
    Smagorinsky τ_eff, clamped to [τ, 10 τ] (cumulant.py:986-997).  The mean is an FP64 device
    reduction read back on the host, as the reference's ``.item()``.
    """
    if f.ndim != 4 or f.shape[0] != 27:
        raise ValueError(f"D3Q27 populations must have shape (27, nz, ny, nx), got {tuple(f.shape)}")
    tau = _tau(tau, "collide_cumulant_d3q27")
    if C_s > 0:
        tau = _domain_mean_smagorinsky_tau(f, tau, C_s, "collide_cumulant_d3q27")
    coll = nat.make_collision(nat.COLL_CUMULANT, tau, eq_rho_floor=False, rates=(omega_b, omega_odd, omega_even))
    return _collide(f, 27, coll, "collide_cumulant_d3q27")


def _domain_mean_smagorinsky_tau(f, tau, C_s, what):
    """Mean over the grid of the per-cell clamped Smagorinsky τ_eff, clamped to [τ, 10 τ] (cumulant.py:986-997)."""
    fc, dims = _pops(f, 27, what)
    acc = torch.zeros(1, dtype=torch.float64, device=fc.device)
    nat.check(nat.lib().tlbm_smagorinsky_tau_sum(27, _eager_dtype(fc, what), fc.data_ptr(), None, *dims, tau,
                                                 float(C_s) ** 2, acc.data_ptr(), nat.stream_of(fc)), what)
    te = float(acc.item()) / (dims[0] * dims[1] * dims[2])
    return max(tau, min(te, 10.0 * tau))


def collide_cumulant_geier_d3q27(f, tau, omega_b=1.0, omega_odd=1.0, omega_even=1.0, C_s=0.0):
    """Geier 2015 cumulant collision (cumulant.py:1050-1156).

    ``C_s > 0``: one rate from the domain-mean Smagorinsky τ_eff, the legacy operator's policy
    (cumulant.py:1101-1110), read back on the host like the reference's ``.item()``.
    """
    if f.ndim != 4 or f.shape[0] != 27:
        raise ValueError(f"D3Q27 populations must have shape (27, nz, ny, nx), got {tuple(f.shape)}")
    tau = _tau(tau, "collide_cumulant_geier_d3q27")
    if C_s > 0.0:
        tau = _domain_mean_smagorinsky_tau(f, tau, C_s, "collide_cumulant_geier_d3q27")
    coll = nat.make_collision(nat.COLL_GEIER, tau, eq_rho_floor=False, rates=(omega_b, omega_odd, omega_even))
    return _collide(f, 27, coll, "collide_cumulant_geier_d3q27")


def _smag(f, q, tau, C_s, what):
    coll = nat.make_collision(nat.COLL_BGK, _tau(tau, what), tau_mode=nat.TAU_SMAG, eq_rho_floor=False,
                              cs_delta_sq=float(C_s) ** 2)
    return _collide(f, q, coll, what)


def collide_smagorinsky_bgk(f, tau, C_s=0.1):
    """D2Q9 BGK with clamped Smagorinsky τ_eff (turbulence.py:216-242)."""
    return _smag(f, 9, tau, C_s, "collide_smagorinsky_bgk")


def collide_smagorinsky_bgk3d(f, tau, C_s=0.1):
    """D3Q19 BGK with clamped Smagorinsky τ_eff (turbulence.py:245-267)."""
    return _smag(f, 19, tau, C_s, "collide_smagorinsky_bgk3d")


def collide_smagorinsky_bgk27(f, tau, C_s=0.1):
    """D3Q27 BGK with clamped Smagorinsky τ_eff (turbulence.py:472-494)."""
    return _smag(f, 27, tau, C_s, "collide_smagorinsky_bgk27")


def collide_smagorinsky_mrt(f, tau, C_s=0.1, s_e=1.64, s_eps=1.54, s_q=1.7):
    """D2Q9 MRT whose stress moments relax at the per-cell Smagorinsky 1/τ_eff (turbulence.py:270-342)."""
    coll = nat.make_collision(nat.COLL_MRT, _tau(tau, "collide_smagorinsky_mrt"), tau_mode=nat.TAU_SMAG,
                              eq_rho_floor=False, cs_delta_sq=float(C_s) ** 2, rates=(s_e, s_eps, s_q, s_e))
    return _collide(f, 9, coll, "collide_smagorinsky_mrt")


def collide_smagorinsky_mrt3d(f, tau, C_s=0.1, s_e=1.19, s_eps=1.4, s_q=1.2, s_pi=None):
    """D3Q19 MRT whose stress moments 9-13 relax at the per-cell Smagorinsky 1/τ_eff (turbulence.py:345-433)."""
    s_pi = s_e if s_pi is None else s_pi
    coll = nat.make_collision(nat.COLL_MRT, _tau(tau, "collide_smagorinsky_mrt3d"), tau_mode=nat.TAU_SMAG,
                              eq_rho_floor=False, cs_delta_sq=float(C_s) ** 2, rates=(s_e, s_eps, s_q, s_pi))
    return _collide(f, 19, coll, "collide_smagorinsky_mrt3d")


def collide_smagorinsky_mrt27(f, tau, C_s=0.1, s_e=1.19, s_eps=1.4, s_q=1.2, s_pi=None):
    """D3Q27 MRT whose stress moments 5-9 relax at the per-cell Smagorinsky 1/τ_eff (turbulence.py:497-587)."""
    s_pi = s_e if s_pi is None else s_pi
    coll = nat.make_collision(nat.COLL_MRT, _tau(tau, "collide_smagorinsky_mrt27"), tau_mode=nat.TAU_SMAG,
                              eq_rho_floor=False, cs_delta_sq=float(C_s) ** 2, rates=(s_e, s_eps, s_q, s_pi))
    return _collide(f, 27, coll, "collide_smagorinsky_mrt27")


# ---------------------------------------------------------------------------
# streaming, bounce-back, reductions
# ---------------------------------------------------------------------------


def _stream(f, q, what):
    f, dims = _pops(f, q, what, contiguous=False)
    store = nat.store_code(f.dtype)
    out = torch.empty(f.shape, dtype=f.dtype, device=f.device)
    rc = nat.lib().tlbm_stream(q, store, f.data_ptr(), nat.strides_of(f), out.data_ptr(), None, *dims,
                               nat.stream_of(f))
    nat.check(rc, what)
    return out


def stream(f):
    """D2Q9 periodic pull stream (solver.py:396-417)."""
    return _stream(f, 9, "stream")


def stream3d(f):
    """D3Q19 periodic pull stream (solver3d.py:210-239); bit-exact with per-lane ``torch.roll``."""
    return _stream(f, 19, "stream3d")


def stream3d_roll(f):
    """Alias of :func:`stream3d` (solver3d.py:266-284); no index tensors are built either way."""
    return _stream(f, 19, "stream3d_roll")


def stream27(f):
    """D3Q27 periodic pull stream (d3q27.py:766-791)."""
    return _stream(f, 27, "stream27")


stream27_roll = stream27


def _bounce(f, mask, q, what):
    f, dims = _pops(f, q, what)
    dt = _eager_dtype(f, what)
    if tuple(mask.shape) != tuple(f.shape[1:]):
        raise ValueError(f"{what}: mask shape {tuple(mask.shape)} must match {tuple(f.shape[1:])}")
    m = mask.to(device=f.device, dtype=torch.int8).contiguous()
    out = torch.empty_like(f)
    rc = nat.lib().tlbm_bounce_back(q, dt, f.data_ptr(), out.data_ptr(), m.data_ptr(),
                                    dims[0] * dims[1] * dims[2], nat.stream_of(f))
    nat.check(rc, what)
    return out


def bounce_back_cells(f, mask):
    """D2Q9 full-way swap ``where(mask, f[opp], f)`` (boundaries.py:39-47)."""
    return _bounce(f, mask, 9, "bounce_back_cells")


def bounce_back_cells_3d(f, mask, f_pre=None):
    """D3Q19 full-way swap (boundaries3d.py:112-132); ``f_pre`` is accepted and ignored, as there."""
    del f_pre
    return _bounce(f, mask, 19, "bounce_back_cells_3d")


def bounce_back_cells_27(f, mask):
    """D3Q27 full-way swap (boundaries_d3q27.py:45-60)."""
    return _bounce(f, mask, 27, "bounce_back_cells_27")


def _mask_arg(mask, f, what):
    if mask is None:
        return None
    if tuple(mask.shape) != tuple(f.shape[1:]):
        raise ValueError(f"{what}: mask shape {tuple(mask.shape)} must match {tuple(f.shape[1:])}")
    return mask.to(device=f.device, dtype=torch.int8).contiguous()


def _boundary(f, bc, what, mask_a=None, mask_b=None):
    f, dims = _pops(f, 19, what)
    dt = _eager_dtype(f, what)
    ma, mb = _mask_arg(mask_a, f, what), _mask_arg(mask_b, f, what)
    out = torch.empty_like(f)
    rc = nat.lib().tlbm_boundary_eager(dt, f.data_ptr(), out.data_ptr(), *dims, bc, nat.ptr(ma), nat.ptr(mb),
                                       nat.stream_of(f))
    nat.check(rc, what)
    return out


def apply_simple_channel_boundaries_3d(f, u_in, wall_mask, obstacle_mask):
    """Equilibrium inlet at x = 0 (u = (u_in, 0, 0), ρ of x = 1), zero-gradient outlet at x = nx−1,
    then bounce-back on ``wall_mask`` and on ``obstacle_mask`` (boundaries3d.py:407-443).

    Returns a new tensor; the input is not modified (the reference writes its inlet and outlet
    planes into ``f`` before bouncing back into a new tensor).
    """
    bc = nat.make_boundary(nat.BC_CHANNEL_X, (float(u_in), 0.0, 0.0))
    return _boundary(f, bc, "apply_simple_channel_boundaries_3d", wall_mask, obstacle_mask)


def far_field_bc_3d(f, u_in, obstacle_mask=None, uy=0.0, uz=0.0, bc_config=None):
    """Free-stream equilibrium (ρ = 1, u = (u_in, uy, uz)) on the inlet and the far-field lateral faces,
    zero-gradient outlet, then bounce-back on ``obstacle_mask`` (boundaries3d.py:446-537).  ``bc_config``
    takes ``far_field_faces`` / ``periodic_faces`` as there; None means all four lateral faces."""
    bc = nat.make_boundary(nat.BC_FAR_FIELD, (float(u_in), float(uy), float(uz)), nat.far_field_faces(bc_config))
    return _boundary(f, bc, "far_field_bc_3d", obstacle_mask)


def zou_he_moving_lid_3d(f, u_lid, uy_lid=0.0, uz_lid=0.0):
    """Zou-He moving lid with u = (u_lid, 0, 0) on every cell of y = ny−1 (boundaries3d.py:324-404)."""
    if uy_lid != 0.0 or uz_lid != 0.0:
        raise NotImplementedError("zou_he_moving_lid_3d supports uy = uz = 0 (the reference's lid-driven cavity case)")
    bc = nat.make_boundary(nat.BC_LID_Y, (float(u_lid), 0.0, 0.0))
    return _boundary(f, bc, "zou_he_moving_lid_3d")


def sphere_mask(nx, ny, nz, cx, cy, cz, radius, device):
    """Boolean (nz, ny, nx) mask of a sphere (boundaries3d.py:29-49), built on the GPU in float32 as there."""
    out = torch.empty((int(nz), int(ny), int(nx)), dtype=torch.bool, device=device)
    nat.require_gpu(out, "sphere_mask")
    rc = nat.lib().tlbm_sphere_mask(int(nx), int(ny), int(nz), float(cx), float(cy), float(cz), float(radius),
                                    out.data_ptr(), nat.stream_of(out))
    nat.check(rc, "sphere_mask")
    return out


def moment_sums(f, *, deviation: bool = False) -> torch.Tensor:
    """FP64 ``(Σf, Σc_x f, Σc_y f, Σc_z f)`` over the whole lattice, one launch, device tensor.

    Any storage dtype and layout; ``deviation=True`` reads 16-bit ``f_i − w_i`` storage.
    """
    q = int(f.shape[0])
    f, dims = _pops(f, q, "moment_sums", contiguous=False)
    out = torch.zeros(4, dtype=torch.float64, device=f.device)
    rc = nat.lib().tlbm_moment_sums(q, nat.store_code(f.dtype, deviation), f.data_ptr(), nat.strides_of(f), *dims,
                                    out.data_ptr(), nat.stream_of(f))
    nat.check(rc, "moment_sums")
    return out


def compute_obstacle_forces_3d(f, obstacle_mask):
    """Ladd momentum-exchange force ``2 Σ_solid Σ_q c_q f`` (obstacles.py:91-130), FP64 accumulated."""
    q = int(f.shape[0])
    f, dims = _pops(f, q, "compute_obstacle_forces_3d", contiguous=False)
    if tuple(obstacle_mask.shape) != tuple(f.shape[1:]):
        raise ValueError("obstacle_mask shape must match the population spatial shape")
    m = obstacle_mask.to(device=f.device, dtype=torch.int8).contiguous()
    acc = torch.zeros(3, dtype=torch.float64, device=f.device)
    rc = nat.lib().tlbm_obstacle_force(q, nat.store_code(f.dtype), f.data_ptr(), nat.strides_of(f), *dims,
                                       m.data_ptr(), acc.data_ptr(), nat.stream_of(f))
    nat.check(rc, "compute_obstacle_forces_3d")
    res = acc.to(f.dtype)
    return res[0], res[1], res[2]


def compute_obstacle_forces_27(f, obstacle_mask):
    """D3Q27 Ladd force ``2 Σ_solid Σ_q c_q f`` (obstacles.py:197-240), FP64 accumulated; the same
    reduction kernel as :func:`compute_obstacle_forces_3d`."""
    if f.ndim != 4 or f.shape[0] != 27:
        raise ValueError(f"compute_obstacle_forces_27: populations must have shape (27, nz, ny, nx), "
                         f"got {tuple(f.shape)}")
    return compute_obstacle_forces_3d(f, obstacle_mask)


def _correct(f, q, target_mass, what):
    f, dims = _pops(f, q, what)
    store = nat.store_code(f.dtype)
    sums = moment_sums(f)
    # scale = target / current, or 1 when the current mass vanishes (solver.py:433-436)
    cur = sums[0:1]
    scale = torch.where(cur.abs() < 1e-30, torch.ones_like(cur), float(target_mass) / cur)
    out = f.clone()
    rc = nat.lib().tlbm_scale(q, store, out.data_ptr(), None, *dims, scale.data_ptr(), nat.stream_of(f))
    nat.check(rc, what)
    return out


def correct_mass(f, target_mass):
    """D2Q9 global mass rescale (solver.py:420-437)."""
    return _correct(f, 9, target_mass, "correct_mass")


def correct_mass3d(f, target_mass):
    """D3Q19 global mass rescale (solver3d.py:287-300), FP64 sum."""
    return _correct(f, 19, target_mass, "correct_mass3d")


def correct_mass27(f, target_mass):
    """D3Q27 global mass rescale (d3q27.py:729-746), FP64 sum."""
    return _correct(f, 27, target_mass, "correct_mass27")


def lattice_of(q: int):
    return BY_Q[q]
