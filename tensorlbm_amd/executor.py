"""``LBMStepExecutor``: the reference's collide → stream → boundary → diagnostics loop on HIP kernels.

Mirror of ``tensorlbm.lbm_step.LBMStepExecutor`` (src/tensorlbm/lbm_step.py:70-686): same
constructor keywords, ``step(f) -> (f, diag)`` and ``run(f, n, *, reporters=None) -> (f, diags)``.

* ``collide_fn="bgk"`` is the one-launch BGK kernel (``f_eq`` built with ρ as in
  ``_collide_bgk``, lbm_step.py:350-376); any callable ``collide_fn(f, tau)`` plugs in, e.g.
  ``tensorlbm_amd.collide_mrt3d``.
* ``stream_fn=None`` is the periodic pull-stream kernel; a callable ``stream_fn(f)`` plugs in.
* ``boundary_fn(f, **boundary_kwargs)`` and ``force_fn(f, rho, ux, uy, uz, **force_kwargs)`` are
  called as there; the macroscopic fields come from the library's moment kernel.
* ``target_mass`` rescales with an FP64 device sum (``tlbm_moment_sums`` + ``tlbm_scale``),
  without a host round trip.
* ``run()`` with the internal BGK and stream and no per-step hooks takes the fused path:
  ``(S∘C)ⁿ = S ∘ (C∘S)ⁿ⁻¹ ∘ C`` — one collide launch, n−1 fused stream+collide launches and one
  stream launch, i.e. one memory pass per step instead of two.

The wall-function stage (``wall_fn=True``) is outside this library's path and raises.
"""

from __future__ import annotations

from typing import Any, Callable, Sequence

import torch

from . import _native as nat
from . import ops
from .reporters import StepContext, dispatch

__all__ = ["LBMStepExecutor"]

_LATTICE_Q = {"D3Q19": 19, "D3Q27": 27}


class LBMStepExecutor:
    def __init__(self, lattice: str, *, collide_fn: str | Callable, stream_fn: Callable | None = None,
                 boundary_fn: Callable | None = None, wall_fn: bool | None = False,
                 force_fn: Callable | None = None, device: torch.device, nx: int, ny: int, nz: int,
                 tau: float = 0.6, dtype: torch.dtype = torch.float32, mask: torch.Tensor | None = None,
                 nu: float = 0.02, y_val: float = 0.5, wall_law: str = "log", target_mass: float | None = None,
                 boundary_kwargs: dict[str, Any] | None = None, force_kwargs: dict[str, Any] | None = None,
                 reporters: Sequence | None = None):
        if lattice not in _LATTICE_Q:
            raise ValueError(f"Unsupported lattice {lattice!r}; supported: {list(_LATTICE_Q)}")
        self._internal_bgk = isinstance(collide_fn, str) and collide_fn == "bgk"
        if not self._internal_bgk and not callable(collide_fn):
            raise ValueError(f"collide_fn must be 'bgk' or a callable, got {collide_fn!r}")
        if wall_fn:
            raise NotImplementedError("the wall-function stage is not part of the HIP lattice-update path")
        self.lattice, self.Q = lattice, _LATTICE_Q[lattice]
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError(f"LBMStepExecutor on the HIP path needs a GPU device, got {self.device}")
        if dtype not in (torch.float32, torch.float64):
            raise ValueError(f"Unsupported dtype {dtype}; use float32 or float64")
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.tau, self.dtype = float(tau), dtype
        self._collide_fn = None if self._internal_bgk else collide_fn
        self._stream_fn = stream_fn
        self._boundary_fn = boundary_fn
        self._force_fn = force_fn
        self._target_mass = target_mass
        self._boundary_kwargs = boundary_kwargs or {}
        self._force_kwargs = force_kwargs or {}
        self.mask = mask
        self.reporters = list(reporters) if reporters else []
        self._report_step = 0
        shape = (self.Q, self.nz, self.ny, self.nx)
        self._a = torch.empty(shape, device=self.device, dtype=dtype)
        self._b = torch.empty(shape, device=self.device, dtype=dtype)
        self._sums = torch.zeros(4, device=self.device, dtype=torch.float64)
        self._coll = nat.make_collision(nat.COLL_BGK, self.tau, eq_rho_floor=False)

    # -- kernels ---------------------------------------------------------------
    def _launch_collide(self, f: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        nat.check(nat.lib().tlbm_collide(self.Q, nat.store_code(self.dtype), f.data_ptr(), None, out.data_ptr(), None,
                                         self.nz, self.ny, self.nx, self._coll, nat.stream_of(f)), "LBMStepExecutor")
        return out

    def _launch_stream(self, f: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        nat.check(nat.lib().tlbm_stream(self.Q, nat.store_code(self.dtype), f.data_ptr(), None, out.data_ptr(), None,
                                        self.nz, self.ny, self.nx, nat.stream_of(f)), "LBMStepExecutor")
        return out

    def _launch_fused(self, f: torch.Tensor, out: torch.Tensor) -> torch.Tensor:
        nat.check(nat.lib().tlbm_step(self.Q, nat.store_code(self.dtype), f.data_ptr(), None, out.data_ptr(), None,
                                      self.nz, self.ny, self.nx, self._coll, None, None, None, nat.stream_of(f)),
                  "LBMStepExecutor")
        return out

    def _macroscopic(self, f: torch.Tensor):
        return (ops.macroscopic3d if self.Q == 19 else ops.macroscopic27)(f)

    def _correct_mass_inplace(self, f: torch.Tensor, target_mass: float) -> None:
        f = f if f.is_contiguous() else f.contiguous()
        self._sums.zero_()
        st = nat.store_code(self.dtype)
        nat.check(nat.lib().tlbm_moment_sums(self.Q, st, f.data_ptr(), None, self.nz, self.ny, self.nx,
                                             self._sums.data_ptr(), nat.stream_of(f)), "correct_mass")
        # scale = target / current where |current| > 1e-30 (lbm_step.py:526-540), on the device
        cur = self._sums[:1]
        scale = torch.where(cur.abs() > 1e-30, float(target_mass) / cur, torch.ones_like(cur))
        nat.check(nat.lib().tlbm_scale(self.Q, st, f.data_ptr(), None, self.nz, self.ny, self.nx, scale.data_ptr(),
                                       nat.stream_of(f)), "correct_mass")

    # -- public API ----------------------------------------------------------------
    def step(self, f: torch.Tensor) -> tuple[torch.Tensor, dict[str, Any]]:
        """One collide → stream → boundary → diagnostics cycle (lbm_step.py:546-613)."""
        diag: dict[str, Any] = {}
        if tuple(f.shape) != (self.Q, self.nz, self.ny, self.nx):
            raise ValueError(f"f must have shape {(self.Q, self.nz, self.ny, self.nx)}, got {tuple(f.shape)}")
        self._a.copy_(f)
        f_post = self._launch_collide(self._a, self._b) if self._internal_bgk else self._collide_fn(self._a, self.tau)
        if self._stream_fn is not None:
            f_s = self._stream_fn(f_post)
        else:
            out = self._a if f_post.data_ptr() != self._a.data_ptr() else torch.empty_like(f_post)
            f_s = self._launch_stream(f_post.contiguous(), out)
        f_bc = self._boundary_fn(f_s, **self._boundary_kwargs) if self._boundary_fn is not None else f_s
        if self._force_fn is not None:
            rho, ux, uy, uz = self._macroscopic(f_bc)
            diag["forces"] = self._force_fn(f_bc, rho, ux, uy, uz, **self._force_kwargs)
            diag["max_speed"] = float(torch.sqrt(ux * ux + uy * uy + uz * uz).max().item())
            diag["mean_rho"] = float(rho.mean().item())
        if self._target_mass is not None:
            if not f_bc.is_contiguous():
                f_bc = f_bc.contiguous()
            self._correct_mass_inplace(f_bc, self._target_mass)
        return f_bc.clone(), diag

    def _fast_path(self) -> bool:
        return (self._internal_bgk and self._stream_fn is None and self._boundary_fn is None
                and self._force_fn is None and self._target_mass is None)

    def run(self, f: torch.Tensor, n_steps: int, *, reporters: Sequence | None = None):
        """``n_steps`` cycles; reporters as in lbm_step.py:615-686."""
        active = self.reporters if reporters is None else list(reporters)
        n_steps = int(n_steps)
        if not active:
            if self._fast_path() and n_steps >= 2:
                f = self._run_fused(f, n_steps)
                self._report_step += n_steps
                return f, [{} for _ in range(n_steps)]
            diags = []
            for _ in range(n_steps):
                f, d = self.step(f)
                diags.append(d)
            self._report_step += n_steps
            return f, diags
        ctx = StepContext(step=self._report_step, f=f, lattice=self.lattice, num_cells=self.nx * self.ny * self.nz,
                          num_steps=n_steps, macroscopic_fn=self._macroscopic)
        step = self._report_step
        diags = []
        for _ in range(n_steps):
            f, d = self.step(f)
            step += 1
            diags.append(d)
            ctx.step, ctx.f, ctx.diag, ctx.num_steps, ctx.stop = step, f, d, n_steps, False
            dispatch(ctx, active)
            if ctx.stop:
                break
        self._report_step = step
        return f, diags

    def _run_fused(self, f: torch.Tensor, n_steps: int) -> torch.Tensor:
        """(S∘C)ⁿ = S ∘ (C∘S)ⁿ⁻¹ ∘ C with the fused pull-stream + collide kernel in the middle."""
        cur = self._launch_collide(f.contiguous(), self._a)
        other = self._b
        for _ in range(n_steps - 1):
            cur, other = self._launch_fused(cur, other), cur
        return self._launch_stream(cur, torch.empty_like(cur))
