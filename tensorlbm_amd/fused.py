"""Fused per-step update: one HIP launch per step.

Drop-in for the reference's fused GPU path:

* :func:`fused_step` / :func:`hip_fused` — periodic pull-stream + BGK
  (``triton_fused``, src/tensorlbm/triton_fused.py:618-667), any lattice,
  fp32 / fp64 / fp16 / bf16 storage, optional ``f_i − w_i`` deviation storage.
* :func:`hip_stream` / :func:`hip_collide` — the separate halves
  (``triton_stream`` / ``triton_collide``, triton_fused.py:526-615).
* :func:`fused_obstacle_step` — pull-stream + full-way bounce-back +
  BGK/CM/CUMULANT with molecular, Smagorinsky or external τ and the fused
  Ladd force (``triton_fused_obstacle_xfar_les``, triton_fused_obstacle.py:1695-2053).
* :class:`FusedSolver3D` — ping-pong solver object (``TritonFusedSolver3D``,
  triton_fused.py:676-962) with ``step``/``run``/``benchmark``.
"""

from __future__ import annotations

import time
from typing import Sequence

import torch

from . import _native as nat
from .precision import PrecisionPolicy

__all__ = [
    "is_available", "fused_step", "hip_fused", "hip_stream", "hip_collide", "fused_obstacle_step",
    "FusedSolver3D", "COLLISIONS", "LATTICES", "padded_populations", "PLANE_PAD",
]

LATTICES = {"D2Q9": nat.D2Q9, "D3Q19": nat.D3Q19, "D3Q27": nat.D3Q27}
COLLISIONS = {"BGK": nat.COLL_BGK, "MRT": nat.COLL_MRT, "CM": nat.COLL_CM, "CUMULANT": nat.COLL_CUMULANT,
              "GEIER": nat.COLL_GEIER}
_DEFAULT_RATES = {
    ("D2Q9", "MRT"): (1.64, 1.54, 1.7, 1.64),
    ("D3Q19", "MRT"): (1.19, 1.4, 1.2, 1.19),
    ("D3Q27", "MRT"): (1.19, 1.4, 1.2, 1.19),
    ("D3Q19", "CM"): (1.0, 1.0, 1.0, 1.0, 1.0),
    ("D3Q27", "CM"): (1.0, 1.0, 1.0, 1.0, 1.0),
    ("D3Q19", "CUMULANT"): (1.0, 1.0, 1.0),
    ("D3Q27", "GEIER"): (1.0, 1.0, 1.0),
}


def is_available() -> bool:
    """True iff a GPU is visible and the HIP library is built."""
    return bool(torch.cuda.is_available()) and nat.library_path().is_file()


# The solver's own buffers: [z][q][y][x] with each q-plane of a z-slab padded
# by PLANE_PAD elements.  Measured on MI355X (D3Q19 fp32, 256^3) this layout
# keeps fewer DRAM rows open per wave than plain [q][z][y][x] SoA.
PLANE_PAD = 64


def padded_populations(q: int, nz: int, ny: int, nx: int, dtype, device, pad: int = PLANE_PAD) -> torch.Tensor:
    """Uninitialised ``(Q, nz, ny, nx)`` view over a [z][q][y][x] buffer with padded q-planes."""
    plane = ny * nx + pad
    base = torch.empty(nz * q * plane, dtype=dtype, device=device)
    view = base.as_strided((q, nz, ny, nx), (plane, q * plane, nx, 1))
    return view if q != 9 else view[:, 0]


def _dims(f: torch.Tensor) -> tuple[int, int, int, int]:
    return nat.grid_of(f)


def _out_for(f: torch.Tensor, out: torch.Tensor | None) -> torch.Tensor:
    if out is None:
        return torch.empty(f.shape, dtype=f.dtype, device=f.device)
    if out.shape != f.shape or out.dtype != f.dtype or out.device != f.device or out.stride(-1) != 1:
        raise ValueError("out must have the shape, dtype and device of f and a unit x stride")
    if out.data_ptr() == f.data_ptr():
        raise ValueError("out aliases f; the pull step reads neighbouring cells (use two buffers)")
    return out


def _launch_step(f, out, coll, solid=None, bc=None, force=None, *, deviation=False, what="fused_step"):
    nat.require_gpu(f, what)
    if f.stride(-1) != 1:
        f = f.contiguous()
    q, nz, ny, nx = _dims(f)
    store = nat.store_code(f.dtype, deviation)
    rc = nat.lib().tlbm_step(q, store, f.data_ptr(), nat.strides_of(f), out.data_ptr(), nat.strides_of(out), nz, ny,
                             nx, coll, nat.ptr(solid), bc, nat.ptr(force), nat.stream_of(f))
    nat.check(rc, what)
    return out


def _launch_step2(f, out, coll, solid=None, bc=None, force=None, *, deviation=False, what="step2"):
    nat.require_gpu(f, what)
    if f.stride(-1) != 1:
        f = f.contiguous()
    q, nz, ny, nx = _dims(f)
    store = nat.store_code(f.dtype, deviation)
    rc = nat.lib().tlbm_step2(q, store, f.data_ptr(), nat.strides_of(f), out.data_ptr(), nat.strides_of(out), nz, ny,
                              nx, coll, nat.ptr(solid), bc, nat.ptr(force), nat.stream_of(f))
    nat.check(rc, what)
    return out


def fused_step(f: torch.Tensor, tau: float, *, out: torch.Tensor | None = None, collision: str = "BGK",
               rates: Sequence[float] | None = None, deviation: bool = False) -> torch.Tensor:
    """Periodic pull-stream + collide in one launch (triton_fused.py:396-436 order: stream, then collide).

    ``deviation=True`` treats 16-bit storage as ``f_i − w_i``.
    """
    out = _out_for(f, out)
    q = int(f.shape[0])
    lat = {9: "D2Q9", 19: "D3Q19", 27: "D3Q27"}.get(q)
    col = collision.upper()
    if col not in COLLISIONS:
        raise ValueError(f"unsupported collision {collision!r}; valid: {', '.join(COLLISIONS)}")
    r = tuple(rates) if rates is not None else _DEFAULT_RATES.get((lat, col), ())
    coll = nat.make_collision(COLLISIONS[col], float(tau), rates=r)
    return _launch_step(f, out, coll, deviation=deviation)


def hip_fused(f: torch.Tensor, tau: float, *, out: torch.Tensor | None = None, **_tiles) -> torch.Tensor:
    """Drop-in for ``triton_fused(f, tau, out=)``; tile arguments are accepted and ignored."""
    return fused_step(f, tau, out=out)


def hip_stream(f: torch.Tensor, *, out: torch.Tensor | None = None, **_tiles) -> torch.Tensor:
    """Periodic pull stream into ``out`` (``triton_stream``); bit-exact with ``torch.roll``."""
    out = _out_for(f, out)
    nat.require_gpu(f, "hip_stream")
    q, nz, ny, nx = _dims(f)
    rc = nat.lib().tlbm_stream(q, nat.store_code(f.dtype), f.data_ptr(), nat.strides_of(f), out.data_ptr(),
                               nat.strides_of(out), nz, ny, nx, nat.stream_of(f))
    nat.check(rc, "hip_stream")
    return out


def hip_collide(f: torch.Tensor, tau: float, *, out: torch.Tensor | None = None, **_tiles) -> torch.Tensor:
    """BGK collide-only with ``ω = 1/τ`` and the floored density in f_eq (``triton_collide``)."""
    nat.require_gpu(f, "hip_collide")
    if out is None:
        out = torch.empty(f.shape, dtype=f.dtype, device=f.device)
    q, nz, ny, nx = _dims(f)
    coll = nat.make_collision(nat.COLL_BGK, float(tau), eq_rho_floor=True)
    rc = nat.lib().tlbm_collide(q, nat.store_code(f.dtype), f.data_ptr(), nat.strides_of(f), out.data_ptr(),
                                nat.strides_of(out), nz, ny, nx, coll, nat.stream_of(f))
    nat.check(rc, "hip_collide")
    return out


def fused_obstacle_step(
    f: torch.Tensor,
    nu_lb: float,
    obstacle: torch.Tensor,
    Cs: float,
    delta: float,
    *,
    collision: str = "BGK",
    out: torch.Tensor | None = None,
    tau_eff: torch.Tensor | None = None,
    fx_buf: torch.Tensor | None = None,
    fy_buf: torch.Tensor | None = None,
    fz_buf: torch.Tensor | None = None,
    wall: str = "fused",
    solid_box: tuple | None = None,
    force: torch.Tensor | None = None,
    **_tiles,
) -> torch.Tensor:
    """Pull-stream + full-way bounce-back + collide + optional Ladd force, one launch.

    Same arguments and semantics as ``triton_fused_obstacle_xfar_les``
    (triton_fused_obstacle.py:1695-2053): τ = 3ν + ½; ω from ``tau_eff``
    (per cell), the clamped Smagorinsky closure (``Cs·delta > 0``) or 1/τ.
    ``wall="split"`` / ``solid_box`` are accepted for compatibility: the
    bounce-back here is a register swap at solid cells, so there is no
    second gather to split off.  Force: when ``fx_buf/fy_buf/fz_buf`` are
    given, the force is ADDED to them (callers zero them, as there); a
    float64 device tensor ``force`` of 3 entries may be passed instead.
    """
    nat.require_gpu(f, "fused_obstacle_step")
    q, nz, ny, nx = _dims(f)
    if q != 19:
        raise ValueError(f"Expected Q=19 (production D3Q19), got Q={q}")
    if tuple(obstacle.shape) != tuple(f.shape[1:]):
        raise ValueError(f"obstacle shape {tuple(obstacle.shape)} does not match f's spatial shape "
                         f"{tuple(f.shape[1:])}")
    if wall not in ("fused", "split"):
        raise ValueError(f"wall must be 'fused' or 'split', got {wall!r}")
    del solid_box
    out = _out_for(f, out)
    tau = 3.0 * float(nu_lb) + 0.5
    col = collision.upper()
    if col not in ("BGK", "CM", "CUMULANT"):
        raise ValueError(f"unsupported collision {collision!r} for the fused obstacle kernel; "
                         "valid: BGK, CM, CUMULANT")
    mode, te = nat.TAU_CONST, None
    if tau_eff is not None:
        if tuple(tau_eff.shape) != (nz, ny, nx):
            raise ValueError(f"tau_eff shape {tuple(tau_eff.shape)} does not match f's spatial shape {(nz, ny, nx)}")
        if tau_eff.dtype != torch.float32:
            raise ValueError(f"tau_eff must be float32, got {tau_eff.dtype}")
        if tau_eff.device != f.device:
            raise ValueError(f"tau_eff device {tau_eff.device} does not match f device {f.device}")
        mode, te = nat.TAU_FIELD, tau_eff.contiguous()
    elif Cs * Cs * delta * delta > 0.0:
        mode = nat.TAU_SMAG
    coll = nat.make_collision(COLLISIONS[col], tau, tau_mode=mode, cs_delta_sq=Cs * Cs * delta * delta,
                              rates=_DEFAULT_RATES.get(("D3Q19", col), ()), tau_eff=te)
    solid = obstacle if obstacle.dtype == torch.int8 else obstacle.to(torch.int8)
    solid = solid.contiguous()
    bufs = (fx_buf, fy_buf, fz_buf)
    acc = None
    if force is not None:
        if force.dtype != torch.float64 or force.numel() != 3 or force.device != f.device:
            raise ValueError("force must be a float64 device tensor with 3 entries")
        acc = force
    elif all(b is not None for b in bufs):
        acc = torch.zeros(3, dtype=torch.float64, device=f.device)
    _launch_step(f, out, coll, solid=solid, force=acc, what="fused_obstacle_step")
    if force is None and acc is not None:
        for i, b in enumerate(bufs):
            b.add_(acc[i].to(b.dtype))
    return out


def apply_far_field_bc_6face(f: torch.Tensor, u_in: float, *, uy: float = 0.0, uz: float = 0.0,
                             feq_pad_buf: torch.Tensor | None = None, feq_vec: torch.Tensor | None = None) -> None:
    """The production companion of the fused obstacle step (triton_fused_obstacle.py:2386-2471), in place:
    free-stream equilibrium (ρ = 1, u = (u_in, uy, uz)) on x = 0, the zero-gradient copy of x = nx − 2 onto
    x = nx − 1, then the equilibrium on y = 0, y = ny − 1 and, when nz > 4 (below that the z axis stays
    the step's periodic wrap, the reference's 2-D-extruded mode), on z = 0 and z = nz − 1, in that order.

    ``f`` holds absolute populations, shape ``(Q, nz, ny, nx)`` with Q ≥ 19 (slots past 19 are set to 0,
    as the reference pads its 32-slot buffers).  ``feq_vec`` (shape ``(19,)`` or ``(Q,)``) skips the
    equilibrium; otherwise it comes from the library's equilibrium kernel on one cell.  These are
    O(Q·area) plane writes on the device, as in the reference; ``FusedSolver3D(boundary="far_field")``
    applies the same faces inside the step instead.
    """
    if f.dim() != 4 or f.shape[0] < 19:
        raise ValueError(f"f must be (Q >= 19, nz, ny, nx), got shape {tuple(f.shape)}")
    nat.require_gpu(f, "apply_far_field_bc_6face")
    q, nz, ny, nx = (int(v) for v in f.shape)
    if feq_vec is None:
        from .ops import equilibrium3d

        one = torch.ones((1, 1, 1), dtype=torch.float64, device=f.device)
        feq_vec = equilibrium3d(one, one * float(u_in), one * float(uy), one * float(uz))[:, 0, 0, 0]
    elif feq_vec.dim() != 1:
        raise ValueError(f"feq_vec must be 1-D shape (Q,), got shape {tuple(feq_vec.shape)}")
    feq_vec = feq_vec.to(device=f.device, dtype=f.dtype)
    if feq_vec.shape[0] < q:
        if feq_pad_buf is None or feq_pad_buf.shape != (q,) or feq_pad_buf.dtype != f.dtype or \
                feq_pad_buf.device != f.device:
            feq_pad_buf = torch.zeros((q,), dtype=f.dtype, device=f.device)
        else:
            feq_pad_buf.zero_()
        feq_pad_buf[: feq_vec.shape[0]] = feq_vec
        feq_vec = feq_pad_buf
    elif feq_vec.shape[0] > q:
        feq_vec = feq_vec[:q]
    face = feq_vec[:, None, None]
    f[:, :, :, 0] = face
    if nx >= 2:
        f[:, :, :, -1] = f[:, :, :, -2]
    f[:, :, 0, :] = face
    if ny >= 2:
        f[:, :, -1, :] = face
    if nz > 4:
        f[:, 0, :, :] = face
        f[:, -1, :, :] = face


def apply_mass_correction(f: torch.Tensor, target_mass: float) -> torch.Tensor:
    """Rescale ``f`` in place so that its sum is ``target_mass`` and return it
    (triton_fused_obstacle.py:2513-2527, the rule of correct_mass3d); the current mass is an FP64 sum on
    the device (the reference sums in the storage type), and a vanishing mass leaves ``f`` as it is."""
    q = int(f.shape[0])
    if f.dim() != 4 or q not in (19, 27):
        raise ValueError(f"f must be (19 or 27, nz, ny, nx), got shape {tuple(f.shape)}")
    nat.require_gpu(f, "apply_mass_correction")
    from .ops import moment_sums

    cur = moment_sums(f)[0:1]
    scale = torch.where(cur.abs() < 1e-30, torch.ones_like(cur), float(target_mass) / cur)
    nz, ny, nx = (int(v) for v in f.shape[1:])
    rc = nat.lib().tlbm_scale(nat.D3Q19 if q == 19 else nat.D3Q27, nat.store_code(f.dtype), f.data_ptr(),
                              nat.strides_of(f), nz, ny, nx, scale.data_ptr(), nat.stream_of(f))
    nat.check(rc, "apply_mass_correction")
    return f


class FusedSolver3D:
    """Ping-pong solver over the fused HIP step (mirror of TritonFusedSolver3D).

    ``step(f)`` returns one of two persistent buffers and never writes into
    its input, so feeding the result back never aliases the pull kernel's
    input and output (triton_fused.py:697-701, 849-855).

    Extra keywords select the lattice (``"D3Q19"`` default, ``"D2Q9"``,
    ``"D3Q27"``), the collision family, deviation storage for 16-bit dtypes,
    an optional solid mask (full-way bounce-back) and a boundary
    (``"channel_x"`` with ``u_bc`` = inlet velocity, ``"lid_y"``, or
    ``"far_field"`` with ``u_bc`` = free stream and far_field_bc_3d's ``bc_config``).
    ``closure=EddyViscosityMLP3D`` relaxes D3Q19 BGK with the learned eddy
    viscosity, evaluated inside the step (``C_s`` then sets the strain's time
    scale: the Smagorinsky τ of that constant, or τ itself for ``C_s=0``).

    ``run()`` advances pairs of steps with one launch each when
    ``temporal_blocking`` is on and the library builds the case: the same
    results as single steps, one HBM round trip of the populations per two
    steps.  ``step2()`` is one such pair.  ``None`` (default) turns it on for
    D3Q19 with fp32 storage, where it pays on MI355X (256³ channel + sphere,
    GLUPS one / two steps per launch: BGK 39.0 / 68.2, MRT 39.6 / 58.2, CM
    39.7 / 55.4, cumulant 39.1 / 59.1); with 16-bit storage (fp16 76.1 / 74.6)
    and for D3Q27 (Geier fp32 27.5 / 28.4, fp16 40.8 / 35.6) the recomputed tile
    halo costs about what the saved traffic gains, or more.
    """

    def __init__(self, nz: int, ny: int, nx: int, tau: float, device: str | torch.device = "cuda:0",
                 dtype: torch.dtype | None = None, *, precision=None, lattice: str = "D3Q19",
                 collision: str = "BGK", rates: Sequence[float] | None = None, deviation: bool = False,
                 C_s: float = 0.0, delta: float = 1.0, solid: torch.Tensor | None = None,
                 boundary: str | None = None, u_bc=(0.0, 0.0, 0.0), layout: str = "padded",
                 closure=None, bc_config: dict | None = None, temporal_blocking: bool | None = None,
                 **_tiles) -> None:
        if precision is not None:
            policy = PrecisionPolicy.parse(precision)
            if policy.compute_dtype != torch.float32:
                raise NotImplementedError(f"{policy.name}: the fused HIP kernel computes in fp32 registers for this "
                                          "tier; use dtype=torch.float64 for an fp64 run")
            if dtype is not None and dtype != policy.store_dtype:
                raise ValueError(f"dtype={dtype} conflicts with precision={policy.name} "
                                 f"(store dtype {policy.store_dtype}); pass only one.")
            dtype = policy.store_dtype
            self.precision_policy = policy
        else:
            self.precision_policy = None
        if dtype is None:
            dtype = torch.float32
        if dtype not in (torch.float32, torch.float64, torch.float16, torch.bfloat16):
            raise ValueError(f"Unsupported dtype {dtype}; use float32, float64, float16, or bfloat16.")
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError(f"FusedSolver3D requires a GPU device, got {self.device}.")
        lat = lattice.upper()
        if lat not in LATTICES:
            raise ValueError(f"unknown lattice {lattice!r}")
        col = collision.upper()
        if col not in COLLISIONS:
            raise ValueError(f"unsupported collision {collision!r}")
        if lat == "D2Q9" and int(nz) != 1:
            raise ValueError("D2Q9 requires nz == 1")
        self.lattice, self.collision = lat, col
        self._Q = LATTICES[lat]
        self.nz, self.ny, self.nx = int(nz), int(ny), int(nx)
        self.tau = float(tau)
        self.dtype = dtype
        self.deviation = bool(deviation)
        self._store = nat.store_code(dtype, self.deviation)
        r = tuple(rates) if rates is not None else _DEFAULT_RATES.get((lat, col), ())
        mode = nat.TAU_SMAG if C_s * C_s * delta * delta > 0.0 else nat.TAU_CONST
        self.closure = closure
        if closure is not None:
            if lat != "D3Q19" or col != "BGK":
                raise NotImplementedError("the learned closure runs in the D3Q19 BGK kernel")
            if not hasattr(closure, "packed"):
                raise TypeError("closure must be an EddyViscosityMLP3D")
            mode = getattr(closure, "tau_mode", nat.TAU_MLP)
        self._coll = nat.make_collision(COLLISIONS[col], self.tau, tau_mode=mode,
                                        cs_delta_sq=C_s * C_s * delta * delta, rates=r)
        self.solid = None
        if solid is not None:
            if tuple(solid.shape) != (self.nz, self.ny, self.nx) and not (lat == "D2Q9" and tuple(solid.shape) == (self.ny, self.nx)):
                raise ValueError("solid mask shape must match the grid")
            self.solid = solid.to(device=self.device, dtype=torch.int8).contiguous()
        kinds = {None: nat.BC_NONE, "channel_x": nat.BC_CHANNEL_X, "lid_y": nat.BC_LID_Y,
                 "far_field": nat.BC_FAR_FIELD}
        if boundary not in kinds:
            raise ValueError(f"unknown boundary {boundary!r}; valid: channel_x, lid_y, far_field")
        faces = nat.far_field_faces(bc_config) if boundary == "far_field" else 0
        self._bc = nat.make_boundary(kinds[boundary], u_bc, faces) if boundary is not None else None
        self.boundary = boundary
        if layout not in ("padded", "contiguous"):
            raise ValueError(f"layout must be 'padded' or 'contiguous', got {layout!r}")
        self.layout = layout
        self._buf_a: torch.Tensor | None = None
        self._buf_b: torch.Tensor | None = None
        self._report_step = 0
        if temporal_blocking is None:
            temporal_blocking = lat == "D3Q19" and self._store == nat.STORE_F32
        self._two_step = bool(temporal_blocking) and lat != "D2Q9" and bool(
            nat.lib().tlbm_step2_supported(self._Q, self._store, self._coll, self._bc, self.nx))

    def _shape(self):
        return (self._Q, self.ny, self.nx) if self.lattice == "D2Q9" else (self._Q, self.nz, self.ny, self.nx)

    @property
    def two_step(self) -> bool:
        """True when ``run()`` / ``step2()`` advance two steps per launch."""
        return self._two_step

    def _check_input(self, f: torch.Tensor) -> None:
        if f.dtype != self.dtype:
            pol = f" (precision {self.precision_policy.name})" if self.precision_policy is not None else ""
            raise ValueError(f"step() input dtype {f.dtype} does not match the solver's storage dtype "
                             f"{self.dtype}{pol}; store the initial distribution in the storage dtype first")
        if tuple(f.shape) != self._shape():
            raise ValueError(f"step() input shape {tuple(f.shape)} != {self._shape()}")

    def _target(self, f: torch.Tensor, force: torch.Tensor | None) -> torch.Tensor:
        if self._buf_a is None:
            self._buf_a = self._alloc()
            self._buf_b = self._alloc()
        if force is not None and self.solid is None:
            raise ValueError("force needs a solid mask")
        return self._buf_b if f.data_ptr() == self._buf_a.data_ptr() else self._buf_a

    def step(self, f: torch.Tensor, *, force: torch.Tensor | None = None) -> torch.Tensor:
        """Advance one step; returns a persistent internal buffer (never ``f`` itself)."""
        self._check_input(f)
        out = self._target(f, force)
        if self.closure is not None:
            self._coll.mlp = self.closure.packed.data_ptr()
        return _launch_step(f, out, self._coll, solid=self.solid, bc=self._bc, force=force,
                            deviation=self.deviation, what="FusedSolver3D.step")

    def step2(self, f: torch.Tensor, *, force: torch.Tensor | None = None) -> torch.Tensor:
        """Advance two steps (one launch when ``two_step``); ``force`` receives both steps' samples.
        D3Q19 samples the Ladd force inside the two-step launch; D3Q27 takes two one-step launches."""
        if not self._two_step or (force is not None and self._Q != 19):
            return self.step(self.step(f, force=force), force=force)
        self._check_input(f)
        out = self._target(f, force)
        return _launch_step2(f, out, self._coll, solid=self.solid, bc=self._bc, force=force,
                             deviation=self.deviation, what="FusedSolver3D.step2")

    def _alloc(self) -> torch.Tensor:
        if self.layout == "contiguous":
            return torch.empty(self._shape(), dtype=self.dtype, device=self.device)
        return padded_populations(self._Q, self.nz, self.ny, self.nx, self.dtype, self.device)

    def run(self, f: torch.Tensor, n_steps: int, *, reporters: Sequence | None = None):
        """Advance ``n_steps``; reporters receive a StepContext after every step (lbm_step.py:615-686)."""
        if not reporters:
            n = int(n_steps)
            if self._two_step:
                for _ in range(n // 2):
                    f = self.step2(f)
                if n % 2:
                    f = self.step(f)
            else:
                for _ in range(n):
                    f = self.step(f)
            self._report_step += n
            return f, [{} for _ in range(int(n_steps))]
        from .reporters import StepContext, dispatch

        ctx = StepContext(step=self._report_step, f=f, lattice=self.lattice,
                          num_cells=self.grid_size_cells(), num_steps=int(n_steps))
        step = ctx.step
        diags: list[dict] = []
        for _ in range(int(n_steps)):
            f = self.step(f)
            step += 1
            diags.append({})
            ctx.step, ctx.f, ctx.num_steps, ctx.stop = step, f, int(n_steps), False
            dispatch(ctx, reporters)
            if ctx.stop:
                break
        self._report_step = step
        return f, diags

    def state_dict(self, f: torch.Tensor, *, step: int | None = None, extra=None) -> dict:
        """Checkpointable state of populations ``f`` (``triton_fused_state_dict``, checkpoint.py:568-586):
        lattice, Q, grid, τ, precision tier, storage dtype and offset convention; ``step`` defaults to the
        ``run()`` counter."""
        from .checkpoint import fused_state_dict

        return fused_state_dict(self, f, step=step, extra=extra)

    def load_state_dict(self, state, *, device=None) -> torch.Tensor:
        """Populations of ``state`` on this solver's device after a fail-closed identity check
        (``triton_fused_load_state_dict``, checkpoint.py:589-617); resets the ``run()`` counter."""
        from .checkpoint import fused_load_state_dict

        return fused_load_state_dict(self, state, device=device)

    def grid_size_cells(self) -> int:
        return self.nz * self.ny * self.nx

    def transient_memory_bytes(self) -> int:
        """Two population buffers (the padded layout adds 64 elements per q-plane per z-slab)."""
        pad = 0 if self.layout == "contiguous" else PLANE_PAD * self._Q * self.nz
        return 2 * (self.grid_size_cells() * self._Q + pad) * self.dtype.itemsize

    def benchmark(self, n_steps: int = 100, warmup: int = 5) -> float:
        """Seconds per step over ``n_steps`` ping-pong steps (wall clock, synchronised)."""
        f = torch.zeros(self._shape(), dtype=self.dtype, device=self.device)
        f += (1e-3 * torch.randn(self._shape(), device=self.device)).to(self.dtype)
        for _ in range(warmup):
            f = self.step(f)
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        for _ in range(n_steps):
            f = self.step(f)
        torch.cuda.synchronize(self.device)
        return (time.perf_counter() - t0) / n_steps
