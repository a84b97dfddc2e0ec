"""Checkpoints of the populations and solver state, file-compatible with the reference.

Two on-disk forms, both readable by the reference package and vice versa:

* the tensor + JSON pair — ``checkpoint_f.pt`` (the populations, CPU copy) and
  ``checkpoint_meta.json`` (``format_version`` 1, ``step``, ``tensor_shape``,
  ``tensor_dtype``, ``lattice_directions`` plus caller extras);
  src/tensorlbm/checkpoint.py:22-140;
* the solver-state envelope — one ``torch.save`` file tagged
  ``"tensorlbm-solver-checkpoint"`` version 2 that holds a state dict
  (``populations``, ``step``, ``lattice``, ``grid``, ``q``, ``tau``, ``dtype``,
  ``obstacle_mask``, ``precision_policy``, adapter extras), caller metadata,
  the torch RNG states and a SHA-256 of every saved tensor's bytes;
  src/tensorlbm/checkpoint.py:181-421.

Loading is fail-closed: a torn or edited file, an unknown version, or a state
whose lattice / Q / grid / tau / precision tier / storage dtype differs from the
solver it is restored into raises :class:`CheckpointError` (a ``ValueError``)
instead of silently changing the trajectory.  Files are read with
``torch.load(weights_only=True)`` only, and written to a temporary name, flushed
to disk and renamed, so an interrupted save leaves the previous file intact.

16-bit deviation storage (``FusedSolver3D(deviation=True)``) keeps ``f_i − w_i``
rather than ``f_i``.  The fused adapters record that as ``storage_offset``
(``"f-w"``; ``"none"`` for plain storage, which is also what a state without
the key means, as every reference state is) and refuse to restore a state into
a solver of the other convention.  :func:`absolute_populations` turns either
kind into plain ``f_i`` in fp32, for handing a deviation state to a solver that
stores ``f_i``.
"""

from __future__ import annotations

import hashlib
import json
import os
from collections.abc import Mapping
from dataclasses import dataclass
from datetime import datetime, timezone
from pathlib import Path
from typing import Any

import torch

__all__ = [
    "CheckpointError", "SolverCheckpoint", "save_checkpoint", "load_checkpoint", "save_solver_checkpoint",
    "load_solver_checkpoint", "eager_state_dict", "eager_load_state_dict", "fused_state_dict",
    "fused_load_state_dict", "triton_fused_state_dict", "triton_fused_load_state_dict", "absolute_populations",
    "atomic_torch_save", "SOLVER_CHECKPOINT_FORMAT", "SOLVER_CHECKPOINT_VERSION", "OFFSET_NONE",
    "OFFSET_WEIGHTS",
]

# file names and format tags shared with the reference (checkpoint.py:22-24, 181-182)
_TENSOR_FILE = "checkpoint_f.pt"
_META_FILE = "checkpoint_meta.json"
_FORMAT_VERSION = 1
SOLVER_CHECKPOINT_FORMAT = "tensorlbm-solver-checkpoint"
SOLVER_CHECKPOINT_VERSION = 2

#: ``storage_offset`` values: populations stored as ``f_i`` / as ``f_i − w_i``
OFFSET_NONE = "none"
OFFSET_WEIGHTS = "f-w"


class CheckpointError(ValueError):
    """A checkpoint that cannot be used as it stands (missing, torn, edited, or for another setup)."""


@dataclass
class SolverCheckpoint:
    """What :func:`load_solver_checkpoint` found, after validation.

    ``state`` goes straight to the matching ``*_load_state_dict``; ``grid`` is
    ``(nz, ny, nx)`` (``(ny, nx)`` in 2-D); ``populations`` is a CPU tensor.
    """

    state: dict[str, Any]
    step: int
    lattice: str | None
    grid: tuple[int, ...] | None
    q: int | None
    metadata: dict[str, Any]
    format_version: int
    path: Path

    @property
    def populations(self) -> torch.Tensor:
        return self.state["populations"]


# --------------------------------------------------------------------------- files


def _write_atomically(path: Path, write) -> Path:
    """Write through ``write(fh)`` into a sibling temporary file, fsync it, then rename over ``path``."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_name(f".{path.name}.tmp-{os.getpid()}")
    try:
        with open(tmp, "wb") as fh:
            write(fh)
            fh.flush()
            os.fsync(fh.fileno())
        os.replace(tmp, path)
    finally:
        if tmp.exists():
            tmp.unlink()
    return path


def atomic_torch_save(payload: Any, path: str | Path) -> Path:
    """``torch.save`` that never leaves a half-written file at ``path`` (checkpoint_io.py:12-24)."""
    return _write_atomically(Path(path), lambda fh: torch.save(payload, fh))


def _cpu_copy(t: torch.Tensor) -> torch.Tensor:
    """Dense CPU copy owning exactly its elements (a view of a padded GPU buffer saves only the view)."""
    return t.detach().to("cpu").clone(memory_format=torch.contiguous_format)


def _is_int(v: object) -> bool:
    return isinstance(v, int) and not isinstance(v, bool)


def save_checkpoint(f: torch.Tensor, step: int, run_dir: str | Path,
                    extra: Mapping[str, object] | None = None) -> Path:
    """Write ``checkpoint_f.pt`` + ``checkpoint_meta.json`` into ``run_dir``; returns ``run_dir``.

    Same files and keys as the reference's ``save_checkpoint`` (checkpoint.py:27-58);
    ``extra`` entries are merged into the JSON and must be JSON-serialisable.
    """
    if not isinstance(f, torch.Tensor):
        raise TypeError(f"f must be a torch.Tensor, got {type(f).__name__}")
    run_dir = Path(run_dir)
    run_dir.mkdir(parents=True, exist_ok=True)
    meta: dict[str, object] = {
        "format_version": _FORMAT_VERSION,
        "step": int(step),
        "tensor_shape": list(f.shape),
        "tensor_dtype": str(f.dtype),
        "lattice_directions": int(f.shape[0]) if f.ndim >= 1 else None,
    }
    if extra:
        meta.update(extra)
    text = (json.dumps(meta, indent=2) + "\n").encode("utf-8")
    atomic_torch_save(_cpu_copy(f), run_dir / _TENSOR_FILE)
    _write_atomically(run_dir / _META_FILE, lambda fh: fh.write(text))
    return run_dir


def load_checkpoint(run_dir: str | Path, device: torch.device | str | None = None, *,
                    expected_shape: tuple[int, ...] | None = None,
                    expected_lattice_directions: int | None = None) -> tuple[torch.Tensor, int, dict[str, object]]:
    """Read the pair written by :func:`save_checkpoint`: ``(f, step, meta)``, ``f`` on ``device`` (CPU default).

    FileNotFoundError when either file is missing; ValueError for metadata that is
    not a JSON object, lacks an integer ``step``, carries a newer
    ``format_version``, or disagrees with the tensor or with ``expected_*``
    (checkpoint.py:61-140).
    """
    run_dir = Path(run_dir)
    tensor_path, meta_path = run_dir / _TENSOR_FILE, run_dir / _META_FILE
    for p, what in ((tensor_path, "tensor"), (meta_path, "metadata")):
        if not p.exists():
            raise FileNotFoundError(f"Checkpoint {what} not found: {p}")
    try:
        meta = json.loads(meta_path.read_text(encoding="utf-8"))
    except json.JSONDecodeError as exc:
        raise ValueError(f"Checkpoint metadata is not valid JSON: {meta_path}") from exc
    if not isinstance(meta, dict) or not all(isinstance(k, str) for k in meta):
        raise ValueError(f"Checkpoint metadata must be a JSON object with string keys: {meta_path}")
    version = meta.get("format_version", 0)
    if not _is_int(version):
        raise ValueError(f"Checkpoint metadata 'format_version' must be an integer: {meta_path}")
    if version > _FORMAT_VERSION:
        raise ValueError(f"Unsupported checkpoint format_version={version}; maximum supported is "
                         f"{_FORMAT_VERSION}: {meta_path}")
    if "step" not in meta:
        raise ValueError(f"Checkpoint metadata missing 'step' key: {meta_path}")
    if not _is_int(meta["step"]):
        raise ValueError(f"Checkpoint metadata 'step' must be an integer: {meta_path}")
    f = torch.load(tensor_path, map_location=torch.device(device) if device is not None else "cpu",
                   weights_only=True)
    if not isinstance(f, torch.Tensor):
        raise ValueError(f"Checkpoint tensor payload is not a torch.Tensor: {tensor_path}")
    shape = tuple(f.shape)
    saved_shape = meta.get("tensor_shape")
    if isinstance(saved_shape, list) and tuple(saved_shape) != shape:
        raise ValueError(f"Checkpoint tensor shape does not match metadata: tensor={shape} "
                         f"metadata={tuple(saved_shape)}")
    if expected_shape is not None and shape != tuple(expected_shape):
        raise ValueError(f"Checkpoint incompatible with current run shape: expected={tuple(expected_shape)} "
                         f"got={shape}")
    if expected_lattice_directions is not None and f.ndim >= 1 and shape[0] != expected_lattice_directions:
        raise ValueError(f"Checkpoint incompatible with current lattice model: expected "
                         f"{expected_lattice_directions} directions, got {shape[0]}")
    return f, int(meta["step"]), meta


# --------------------------------------------------------------------------- solver-state envelope


def _digest(t: torch.Tensor) -> str:
    """SHA-256 of a tensor's bytes in C order (the reference's integrity record, checkpoint.py:265-268)."""
    flat = t.detach().to("cpu").contiguous().reshape(-1)
    return hashlib.sha256(flat.view(torch.uint8).numpy()).hexdigest()


def _require_state(state: object) -> dict[str, Any]:
    """The keys every state needs, with their types (checkpoint.py:232-262)."""
    if not isinstance(state, Mapping):
        raise CheckpointError(f"state must be a dict, got {type(state).__name__}")
    state = dict(state)
    missing = [k for k in ("populations", "step") if k not in state]
    if missing:
        raise CheckpointError(f"state dict is missing required key(s): {', '.join(missing)}")
    f = state["populations"]
    if not isinstance(f, torch.Tensor):
        raise CheckpointError(f"state['populations'] must be a torch.Tensor, got {type(f).__name__}")
    if f.ndim < 1:
        raise CheckpointError("state['populations'] needs at least the lattice dimension")
    if not _is_int(state["step"]) or state["step"] < 0:
        raise CheckpointError(f"state['step'] must be a non-negative integer, got {state['step']!r}")
    lattice = state.get("lattice")
    if lattice is not None and not isinstance(lattice, str):
        raise CheckpointError(f"state['lattice'] must be a string or None, got {lattice!r}")
    grid = state.get("grid")
    if grid is not None and not (isinstance(grid, (tuple, list)) and all(_is_int(n) and n > 0 for n in grid)):
        raise CheckpointError(f"state['grid'] must be a tuple of positive ints, got {grid!r}")
    q = state.get("q")
    if q is not None and not (_is_int(q) and q > 0):
        raise CheckpointError(f"state['q'] must be a positive integer or None, got {q!r}")
    return state


def _rng_snapshot(like: torch.Tensor) -> dict[str, Any]:
    rng: dict[str, Any] = {"cpu": torch.get_rng_state()}
    if like.device.type == "cuda":
        rng["cuda"] = torch.cuda.get_rng_state(like.device)
        rng["cuda_device"] = str(like.device)
    return rng


def _rng_restore(rng: object) -> None:
    if not isinstance(rng, dict) or not isinstance(rng.get("cpu"), torch.Tensor):
        return
    torch.set_rng_state(rng["cpu"])
    gpu = rng.get("cuda")
    if not isinstance(gpu, torch.Tensor):
        return
    dev = rng.get("cuda_device")
    index = torch.device(dev).index if isinstance(dev, str) else None
    if not torch.cuda.is_available() or (index is not None and index >= torch.cuda.device_count()):
        raise CheckpointError(f"the checkpoint's RNG state belongs to GPU {dev or 'current'}, which this host "
                              "does not have; load with restore_rng=False to skip it (a continuation that draws "
                              "random numbers is then not bit-identical)")
    torch.cuda.set_rng_state(gpu, dev)


def save_solver_checkpoint(path: str | Path, state: Mapping[str, Any], *,
                           metadata: Mapping[str, Any] | None = None) -> Path:
    """Write a state dict (from an adapter) as one tagged, digested envelope; returns the path.

    Tensors are stored as CPU copies, so the file loads on any device
    (checkpoint.py:296-338).
    """
    payload = _require_state(state)
    digests: dict[str, str] = {}
    for key in ("populations", "obstacle_mask"):
        value = payload.get(key)
        if isinstance(value, torch.Tensor):
            payload[key] = _cpu_copy(value)
            digests[key] = _digest(payload[key])
    grid = payload.get("grid")
    envelope = {
        "format": SOLVER_CHECKPOINT_FORMAT,
        "format_version": SOLVER_CHECKPOINT_VERSION,
        "created_at": datetime.now(timezone.utc).isoformat(),
        "step": int(payload["step"]),
        "lattice": payload.get("lattice"),
        "grid": tuple(grid) if grid is not None else None,
        "q": payload.get("q"),
        "state": payload,
        "metadata": dict(metadata or {}),
        "rng": _rng_snapshot(state["populations"]),
        "digests": digests,
    }
    return atomic_torch_save(envelope, path)


def load_solver_checkpoint(path: str | Path, *, expected_lattice: str | None = None,
                           expected_grid: tuple[int, ...] | None = None, expected_q: int | None = None,
                           restore_rng: bool = True) -> SolverCheckpoint:
    """Read and check an envelope written by :func:`save_solver_checkpoint` (checkpoint.py:341-421).

    CheckpointError for a missing file, one that does not decode, lacks the
    format tag, has another version, fails a digest, or differs from an
    ``expected_*`` identity.  ``restore_rng`` puts the saved torch RNG states back.
    """
    target = Path(path)
    if not target.is_file():
        raise CheckpointError(f"checkpoint file not found: {target}")
    try:
        envelope = torch.load(target, map_location="cpu", weights_only=True)
    except Exception as exc:  # any decode failure means the file is unusable
        raise CheckpointError(f"checkpoint {target} could not be decoded (truncated or corrupt write?): "
                              f"{type(exc).__name__}: {exc}") from exc
    if not isinstance(envelope, dict) or envelope.get("format") != SOLVER_CHECKPOINT_FORMAT:
        raise CheckpointError(f"{target} is not a solver checkpoint (missing format tag "
                              f"{SOLVER_CHECKPOINT_FORMAT!r})")
    version = envelope.get("format_version")
    if version != SOLVER_CHECKPOINT_VERSION:
        raise CheckpointError(f"unsupported checkpoint format_version={version!r} in {target}; "
                              f"this reader supports {SOLVER_CHECKPOINT_VERSION}")
    state = _require_state(envelope.get("state"))
    digests = envelope.get("digests")
    if isinstance(digests, dict):
        for key, want in digests.items():
            value = state.get(key)
            if not isinstance(value, torch.Tensor) or not isinstance(want, str):
                raise CheckpointError(f"checkpoint {target} has an invalid integrity record for {key!r}")
            if _digest(value) != want:
                raise CheckpointError(f"checkpoint {target} failed the integrity digest for {key!r} "
                                      "(payload differs from what was saved)")
    lattice, q = state.get("lattice"), state.get("q")
    grid = tuple(state["grid"]) if state.get("grid") is not None else None
    if expected_lattice is not None and lattice != expected_lattice:
        raise CheckpointError(f"checkpoint lattice mismatch: expected {expected_lattice!r}, got {lattice!r}")
    if expected_grid is not None and grid != tuple(expected_grid):
        raise CheckpointError(f"checkpoint grid mismatch: expected {tuple(expected_grid)}, got {grid}")
    if expected_q is not None and q != expected_q:
        raise CheckpointError(f"checkpoint Q mismatch: expected {expected_q}, got {q}")
    if restore_rng:
        _rng_restore(envelope.get("rng"))
    return SolverCheckpoint(state=state, step=int(state["step"]), lattice=lattice if isinstance(lattice, str) else None,
                            grid=grid, q=q if _is_int(q) else None, metadata=dict(envelope.get("metadata") or {}),
                            format_version=int(version), path=target)


# --------------------------------------------------------------------------- solver adapters


def _solver_grid(solver: Any, f: torch.Tensor | None = None) -> tuple[int, ...] | None:
    """The solver's spatial grid: ``(ny, nx)`` for D2Q9, else ``(nz, ny, nx)`` when all three are ints."""
    dims = [getattr(solver, n, None) for n in ("nz", "ny", "nx")]
    if not all(_is_int(d) for d in dims):
        return tuple(int(n) for n in f.shape[1:]) if f is not None else None
    if str(getattr(solver, "lattice", "")).upper() == "D2Q9":
        return (int(dims[1]), int(dims[2]))
    return tuple(int(d) for d in dims)


def _solver_q(solver: Any) -> int | None:
    for name in ("Q", "_Q"):
        v = getattr(solver, name, None)
        if _is_int(v):
            return int(v)
    return None


def _solver_mask(solver: Any) -> torch.Tensor | None:
    for name in ("mask", "_mask", "solid_mask", "_solid_mask", "solid"):
        v = getattr(solver, name, None)
        if isinstance(v, torch.Tensor):
            return v
    return None


def _policy_name(solver: Any) -> str | None:
    return getattr(getattr(solver, "precision_policy", None), "name", None)


def _storage_offset(solver: Any) -> str:
    return OFFSET_WEIGHTS if getattr(solver, "deviation", False) else OFFSET_NONE


def _base_state(solver: Any, f: torch.Tensor, step: int | None, extra: Mapping[str, Any] | None) -> dict[str, Any]:
    """Fields common to every adapter (checkpoint.py:466-494)."""
    if not isinstance(f, torch.Tensor):
        raise TypeError(f"populations must be a torch.Tensor, got {type(f).__name__}")
    tau = getattr(solver, "tau", None)
    q = _solver_q(solver)
    state: dict[str, Any] = {
        "populations": f.detach(),
        "step": int(step) if step is not None else int(getattr(solver, "_report_step", 0) or 0),
        "lattice": getattr(solver, "lattice", None),
        "grid": _solver_grid(solver, f),
        "q": q if q is not None else int(f.shape[0]),
        "tau": float(tau) if tau is not None else None,
        "dtype": str(f.dtype),
        "obstacle_mask": _solver_mask(solver),
        "precision_policy": _policy_name(solver),
    }
    if extra:
        state["extra"] = dict(extra)
    return state


def _restore(solver: Any, state: Mapping[str, Any], device) -> torch.Tensor:
    """Check a state against the solver's identity and return its populations on the solver's device."""
    s = _require_state(state)
    f = s["populations"]
    lattice, mine = s.get("lattice"), getattr(solver, "lattice", None)
    if isinstance(lattice, str) and isinstance(mine, str) and lattice.upper() != mine.upper():
        raise CheckpointError(f"state lattice {lattice!r} does not match solver lattice {mine!r}")
    q = _solver_q(solver)
    saved_q = s.get("q") if s.get("q") is not None else int(f.shape[0])
    if q is not None and (saved_q != q or int(f.shape[0]) != q):
        raise CheckpointError(f"state Q {saved_q} (populations {int(f.shape[0])}) does not match solver Q {q}")
    grid = _solver_grid(solver)
    if grid is not None:
        saved = tuple(s.get("grid") or ())
        if saved != grid or tuple(f.shape[1:]) != grid:
            raise CheckpointError(f"state grid {saved} (populations {tuple(f.shape[1:])}) does not match "
                                  f"solver grid {grid}")
    tau, my_tau = s.get("tau"), getattr(solver, "tau", None)
    if tau is not None and my_tau is not None and float(tau) != float(my_tau):
        raise CheckpointError(f"state tau {tau} does not match solver tau {my_tau}")
    target = device if device is not None else getattr(solver, "device", None)
    return f.to(device=target) if target is not None else f


def _restore_step(solver: Any, state: Mapping[str, Any]) -> None:
    if _is_int(getattr(solver, "_report_step", None)):
        solver._report_step = int(state["step"])


def eager_state_dict(solver: Any, f: torch.Tensor, *, step: int | None = None,
                     extra: Mapping[str, Any] | None = None) -> dict[str, Any]:
    """State dict of an eager solver (``LBMStepExecutor`` or anything with ``tau`` and grid attributes).

    ``step`` defaults to the solver's ``run()`` counter (checkpoint.py:520-547).
    """
    return _base_state(solver, f, step, extra)


def eager_load_state_dict(solver: Any, state: Mapping[str, Any], *,
                          device: torch.device | str | None = None) -> torch.Tensor:
    """Check lattice / Q / grid / tau, reset the solver's step counter, return the populations (checkpoint.py:550-565)."""
    f = _restore(solver, state, device)
    _restore_step(solver, state)
    return f


def fused_state_dict(solver: Any, f: torch.Tensor, *, step: int | None = None,
                     extra: Mapping[str, Any] | None = None) -> dict[str, Any]:
    """State dict of a :class:`~tensorlbm_amd.FusedSolver3D` (``triton_fused_state_dict``, checkpoint.py:568-586).

    Adds ``storage_dtype`` and the solver identity that shapes the stored bits:
    ``storage_offset`` (``"f-w"`` under deviation storage) and, where the solver
    has them, ``collision`` and ``boundary``.
    """
    state = _base_state(solver, f, step, extra)
    if state["lattice"] is None:
        state["lattice"] = "D3Q19"
    state["storage_dtype"] = str(getattr(solver, "dtype", f.dtype))
    state["storage_offset"] = _storage_offset(solver)
    for key in ("collision", "boundary"):
        value = getattr(solver, key, None)
        if isinstance(value, str):
            state[key] = value
    return state


def fused_load_state_dict(solver: Any, state: Mapping[str, Any], *,
                          device: torch.device | str | None = None) -> torch.Tensor:
    """Restore onto a :class:`~tensorlbm_amd.FusedSolver3D`, fail-closed (checkpoint.py:589-617).

    Besides lattice / Q / grid / tau: the precision tier, the storage dtype (of
    the record and of the tensor), the storage offset convention, and the
    collision family and boundary where both sides name them.
    """
    saved_policy, policy = state.get("precision_policy"), _policy_name(solver)
    if saved_policy != policy:
        raise CheckpointError(f"state precision policy {saved_policy!r} does not match solver policy {policy!r}")
    saved_dtype = state.get("storage_dtype") or state.get("dtype")
    dtype = getattr(solver, "dtype", None)
    if dtype is not None and str(saved_dtype) != str(dtype):
        raise CheckpointError(f"state storage dtype {saved_dtype} does not match solver dtype {dtype}")
    f = state.get("populations")
    if dtype is not None and isinstance(f, torch.Tensor) and f.dtype != dtype:
        raise CheckpointError(f"state populations are {f.dtype} but the record says {saved_dtype}")
    saved_offset = state.get("storage_offset") or OFFSET_NONE
    if saved_offset != _storage_offset(solver):
        raise CheckpointError(f"state storage offset {saved_offset!r} does not match the solver's "
                              f"{_storage_offset(solver)!r} (deviation storage keeps f - w); convert with "
                              "absolute_populations() or build the solver with the matching deviation flag")
    for key in ("collision", "boundary"):
        saved, mine = state.get(key), getattr(solver, key, None)
        if isinstance(saved, str) and isinstance(mine, str) and saved.upper() != mine.upper():
            raise CheckpointError(f"state {key} {saved!r} does not match solver {key} {mine!r}")
    f = _restore(solver, state, device)
    _restore_step(solver, state)
    return f


# the reference's adapter names (checkpoint.py:568, 589)
triton_fused_state_dict = fused_state_dict
triton_fused_load_state_dict = fused_load_state_dict


def absolute_populations(state: Mapping[str, Any]) -> torch.Tensor:
    """The state's populations as plain ``f_i`` in fp32 (adds ``w_i`` back under ``"f-w"`` storage).

    The fused kernels widen a stored deviation to fp32 and add the fp32 weight,
    so this is the value the next step starts from.
    """
    s = _require_state(state)
    f = s["populations"].float()
    offset = s.get("storage_offset") or OFFSET_NONE
    if offset == OFFSET_NONE:
        return f
    if offset != OFFSET_WEIGHTS:
        raise CheckpointError(f"unknown storage offset {offset!r}")
    from .lattice import BY_Q

    q = int(f.shape[0])
    if q not in BY_Q:
        raise CheckpointError(f"no lattice with {q} directions for the {offset!r} offset")
    w = BY_Q[q].W.to(f.device)
    return f + w.reshape((q,) + (1,) * (f.ndim - 1))
