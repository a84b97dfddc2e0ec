"""Checkpoints of the populations and of a solver's state, in the reference's file formats.

Two on-disk forms are written and read; their file names, keys, format tag,
version numbers and digest are the reference's (src/tensorlbm/checkpoint.py),
so either package resumes the other's files:

* a run directory holding ``checkpoint_f.pt`` (the populations as a dense
  CPU tensor) and ``checkpoint_meta.json`` (``format_version`` 1, ``step``,
  ``tensor_shape``, ``tensor_dtype``, ``lattice_directions`` and any caller
  keys);
* one ``torch.save`` envelope tagged ``"tensorlbm-solver-checkpoint"``,
  version 2: a state dict, caller metadata, torch's RNG states and a SHA-256
  of the bytes of each saved tensor.

Design.  A state dict is checked against a declarative schema
(``_STATE_SCHEMA``).  What a continuation depends on -- lattice, Q, grid, τ,
precision tier, stored format, collision and boundary -- is gathered into a
:class:`RunIdentity`, once from the solver and once from the state, and the
two are compared field by field; every disagreement raises
:class:`CheckpointError` (a ``ValueError``).  Files are read with
``torch.load(weights_only=True)`` only and written through a temporary file
in the same directory that is flushed, fsynced and renamed.

Stored format.  16-bit deviation storage keeps ``f_i − w_i``.  Such a state
records ``storage_dtype`` as ``"<dtype>:f-w"`` (e.g. ``"torch.float16:f-w"``)
and ``storage_offset`` ``"f-w"``.  The reference's fused loader compares
``storage_dtype`` with its solver's dtype, so it refuses such a state rather
than reading ``f − w`` as ``f``; a plain state (``"torch.float16"``, offset
``"none"``) is exactly what the reference writes.  :func:`absolute_populations`
turns either kind into plain ``f_i`` in fp32.
"""

from __future__ import annotations

import hashlib
import json
import os
import tempfile
from collections.abc import Callable, Mapping
from dataclasses import dataclass
from datetime import datetime, timezone
from pathlib import Path
from typing import Any, BinaryIO

import torch

__all__ = [
    "CheckpointError", "SolverCheckpoint", "RunIdentity", "save_checkpoint", "load_checkpoint",
    "save_solver_checkpoint", "load_solver_checkpoint", "eager_state_dict", "eager_load_state_dict",
    "fused_state_dict", "fused_load_state_dict", "triton_fused_state_dict", "triton_fused_load_state_dict",
    "absolute_populations", "atomic_torch_save", "SOLVER_CHECKPOINT_FORMAT", "SOLVER_CHECKPOINT_VERSION",
    "OFFSET_NONE", "OFFSET_WEIGHTS",
]

# ------------------------------------------------------------------ the shared on-disk contract

PAIR_TENSOR_FILE = "checkpoint_f.pt"
PAIR_META_FILE = "checkpoint_meta.json"
PAIR_FORMAT_VERSION = 1
SOLVER_CHECKPOINT_FORMAT = "tensorlbm-solver-checkpoint"
SOLVER_CHECKPOINT_VERSION = 2
DIGESTED_KEYS = ("populations", "obstacle_mask")

#: stored-format conventions: ``f_i`` itself / ``f_i − w_i``
OFFSET_NONE = "none"
OFFSET_WEIGHTS = "f-w"
_OFFSETS = (OFFSET_NONE, OFFSET_WEIGHTS)


class CheckpointError(ValueError):
    """A checkpoint that cannot be used as it stands: missing, torn, edited, or for another run."""


def _plain_int(v: object) -> bool:
    return isinstance(v, int) and not isinstance(v, bool)


# ------------------------------------------------------------------ durable writes


def _replace_durably(path: Path, write: Callable[[BinaryIO], object]) -> Path:
    """Run ``write`` on a fresh temporary file beside ``path``, make it durable, then rename it over ``path``.

    A crash at any point leaves either the old file or the complete new one.
    """
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    fd, tmp_name = tempfile.mkstemp(prefix=f".{path.name}.", suffix=".partial", dir=path.parent)
    tmp = Path(tmp_name)
    try:
        with os.fdopen(fd, "wb") as fh:
            write(fh)
            fh.flush()
            os.fsync(fh.fileno())
        os.replace(tmp, path)
        try:  # make the rename itself durable where the platform allows it
            dfd = os.open(path.parent, os.O_RDONLY)
        except OSError:
            dfd = None
        if dfd is not None:
            try:
                os.fsync(dfd)
            except OSError:
                pass
            finally:
                os.close(dfd)
    finally:
        if tmp.exists():
            tmp.unlink()
    return path


def atomic_torch_save(payload: Any, path: str | Path) -> Path:
    """``torch.save`` that never leaves a partly written file at ``path`` (checkpoint_io.py:12-24)."""
    return _replace_durably(Path(path), lambda fh: torch.save(payload, fh))


def _dense_cpu(t: torch.Tensor) -> torch.Tensor:
    """A CPU tensor owning exactly ``t``'s elements, C-ordered (a strided GPU view saves only the view)."""
    return t.detach().to("cpu").contiguous().clone()


def _sha256(t: torch.Tensor) -> str:
    """SHA-256 over the tensor's bytes in C order (the digest both packages record)."""
    raw = _dense_cpu(t).reshape(-1).view(torch.uint8).numpy()
    return hashlib.sha256(raw.tobytes()).hexdigest()


# ------------------------------------------------------------------ the tensor + JSON pair


def save_checkpoint(f: torch.Tensor, step: int, run_dir: str | Path,
                    extra: Mapping[str, object] | None = None) -> Path:
    """Write ``checkpoint_f.pt`` and ``checkpoint_meta.json`` into ``run_dir`` and return ``run_dir``.

    The reference's ``save_checkpoint`` (checkpoint.py:27-58) writes the same files and keys;
    ``extra`` entries join the JSON object and must be JSON-serialisable.
    """
    if not isinstance(f, torch.Tensor):
        raise TypeError(f"f must be a torch.Tensor, got {type(f).__name__}")
    run_dir = Path(run_dir)
    record: dict[str, object] = dict(
        format_version=PAIR_FORMAT_VERSION,
        step=int(step),
        tensor_shape=[int(n) for n in f.shape],
        tensor_dtype=str(f.dtype),
        lattice_directions=int(f.shape[0]) if f.ndim else None,
    )
    record.update(extra or {})
    text = json.dumps(record, indent=2) + "\n"
    atomic_torch_save(_dense_cpu(f), run_dir / PAIR_TENSOR_FILE)
    _replace_durably(run_dir / PAIR_META_FILE, lambda fh: fh.write(text.encode("utf-8")))
    return run_dir


def _pair_record(path: Path) -> dict:
    """The JSON half of a pair, with its ``format_version`` and ``step`` checked."""
    try:
        record = json.loads(path.read_text(encoding="utf-8"))
    except json.JSONDecodeError as exc:
        raise ValueError(f"{path}: the checkpoint metadata is not valid JSON") from exc
    if not isinstance(record, dict):
        raise ValueError(f"{path}: the checkpoint metadata must be a JSON object")
    version = record.get("format_version", 0)  # files from before versioning carry none
    if not _plain_int(version):
        raise ValueError(f"{path}: 'format_version' must be an integer, got {version!r}")
    if version > PAIR_FORMAT_VERSION:
        raise ValueError(f"Unsupported checkpoint format_version={version} in {path} "
                         f"(this reader knows up to {PAIR_FORMAT_VERSION})")
    if "step" not in record:
        raise ValueError(f"{path}: checkpoint metadata missing 'step'")
    if not _plain_int(record["step"]):
        raise ValueError(f"{path}: 'step' must be an integer, got {record['step']!r}")
    return record


def load_checkpoint(run_dir: str | Path, device: torch.device | str | None = None, *,
                    expected_shape: tuple[int, ...] | None = None,
                    expected_lattice_directions: int | None = None) -> tuple[torch.Tensor, int, dict[str, object]]:
    """Read a pair written by :func:`save_checkpoint`: ``(f, step, meta)`` with ``f`` on ``device`` (CPU default).

    FileNotFoundError when a file is missing; ValueError when the metadata is
    malformed, newer than this reader, or disagrees with the tensor or with
    ``expected_shape`` / ``expected_lattice_directions`` (checkpoint.py:61-140).
    """
    run_dir = Path(run_dir)
    tensor_path, meta_path = run_dir / PAIR_TENSOR_FILE, run_dir / PAIR_META_FILE
    missing = [p for p in (tensor_path, meta_path) if not p.exists()]
    if missing:
        raise FileNotFoundError(f"checkpoint file not found: {missing[0]}")
    record = _pair_record(meta_path)
    f = torch.load(tensor_path, map_location=torch.device(device) if device is not None else "cpu",
                   weights_only=True)
    if not isinstance(f, torch.Tensor):
        raise ValueError(f"{tensor_path}: the checkpoint payload is not a torch.Tensor")
    shape = tuple(f.shape)
    recorded = record.get("tensor_shape")
    if isinstance(recorded, list) and tuple(recorded) != shape:
        raise ValueError(f"{tensor_path}: tensor of shape {shape} does not match metadata shape {tuple(recorded)}")
    if expected_shape is not None and shape != tuple(expected_shape):
        raise ValueError(f"checkpoint shape {shape} does not fit the current run shape {tuple(expected_shape)}")
    if expected_lattice_directions is not None and f.ndim and shape[0] != expected_lattice_directions:
        raise ValueError(f"checkpoint lattice: expected {expected_lattice_directions} directions, got {shape[0]}")
    return f, int(record["step"]), record


# ------------------------------------------------------------------ state dicts: schema and identity


def _is_grid(v: object) -> bool:
    return isinstance(v, (tuple, list)) and all(_plain_int(n) and n > 0 for n in v)


# (key, required, predicate, message when the predicate fails); checked in this order
_STATE_SCHEMA: tuple[tuple[str, bool, Callable[[object], bool], str], ...] = (
    ("populations", True, lambda v: isinstance(v, torch.Tensor), "must be a torch.Tensor"),
    ("populations", True, lambda v: v.ndim >= 1, "needs at least the lattice dimension"),
    ("step", True, lambda v: _plain_int(v) and v >= 0, "must be a non-negative integer"),
    ("lattice", False, lambda v: v is None or isinstance(v, str), "must be a string or None"),
    ("grid", False, lambda v: v is None or _is_grid(v), "must be a tuple of positive ints"),
    ("q", False, lambda v: v is None or (_plain_int(v) and v > 0), "must be a positive integer or None"),
)


def _checked_state(state: object) -> dict[str, Any]:
    """A shallow copy of ``state`` after the schema's checks."""
    if not isinstance(state, Mapping):
        raise CheckpointError(f"state must be a dict, got {type(state).__name__}")
    out = dict(state)
    absent = sorted({k for k, required, _, _ in _STATE_SCHEMA if required and k not in out})
    if absent:
        raise CheckpointError(f"state dict is missing required key(s): {', '.join(absent)}")
    for key, _, ok, message in _STATE_SCHEMA:
        if key in out and not ok(out[key]):
            value = out[key]
            shown = type(value).__name__ if key == "populations" else repr(value)
            raise CheckpointError(f"state['{key}'] {message} (got {shown})")
    return out


def _encode_storage(dtype: object, offset: str) -> str:
    return str(dtype) if offset == OFFSET_NONE else f"{dtype}:{offset}"


def _decode_storage(text: object, offset_key: object) -> tuple[str | None, str]:
    """(dtype string, offset) of a state's stored format; a bare dtype means ``f_i`` itself."""
    if offset_key is not None and offset_key not in _OFFSETS:
        raise CheckpointError(f"unknown storage offset {offset_key!r} (known: {', '.join(_OFFSETS)})")
    dtype, offset = (None, None) if text is None else (str(text), None)
    if dtype is not None and ":" in dtype:
        dtype, offset = dtype.split(":", 1)
    if offset_key is not None and offset is not None and offset_key != offset:
        raise CheckpointError(f"storage_dtype names offset {offset!r} but storage_offset is {offset_key!r}")
    offset = offset or (offset_key if offset_key is not None else OFFSET_NONE)
    if offset not in _OFFSETS:
        raise CheckpointError(f"unknown storage offset {offset!r} (known: {', '.join(_OFFSETS)})")
    return dtype, offset


@dataclass(frozen=True)
class RunIdentity:
    """What a continuation of a run depends on; ``None`` means "not recorded / not known"."""

    lattice: str | None = None
    q: int | None = None
    grid: tuple[int, ...] | None = None
    tau: float | None = None
    precision_policy: str | None = None
    storage_dtype: str | None = None
    storage_offset: str = OFFSET_NONE
    collision: str | None = None
    boundary: str | None = None

    @staticmethod
    def of_solver(solver: Any, f: torch.Tensor | None = None) -> "RunIdentity":
        """Read from a solver's attributes (``FusedSolver3D``, ``LBMStepExecutor`` or a look-alike)."""
        lattice = getattr(solver, "lattice", None)
        q = next((v for v in (getattr(solver, n, None) for n in ("Q", "q", "_Q")) if _plain_int(v)), None)
        dims = tuple(getattr(solver, n, None) for n in ("nz", "ny", "nx"))
        if all(_plain_int(d) for d in dims):
            grid = dims[1:] if isinstance(lattice, str) and lattice.upper() == "D2Q9" else dims
        else:
            grid = tuple(int(n) for n in f.shape[1:]) if f is not None else None
        tau = getattr(solver, "tau", None)
        dtype = getattr(solver, "dtype", None)
        text = lambda v: v if isinstance(v, str) else None  # noqa: E731
        return RunIdentity(
            lattice=text(lattice), q=q, grid=tuple(int(n) for n in grid) if grid is not None else None,
            tau=float(tau) if tau is not None else None,
            precision_policy=getattr(getattr(solver, "precision_policy", None), "name", None),
            storage_dtype=str(dtype) if dtype is not None else None,
            storage_offset=OFFSET_WEIGHTS if getattr(solver, "deviation", False) else OFFSET_NONE,
            collision=text(getattr(solver, "collision", None)), boundary=text(getattr(solver, "boundary", None)))

    @staticmethod
    def of_state(state: Mapping[str, Any]) -> "RunIdentity":
        f = state["populations"]
        dtype, offset = _decode_storage(state.get("storage_dtype") or state.get("dtype"), state.get("storage_offset"))
        grid = state.get("grid")
        tau = state.get("tau")
        return RunIdentity(
            lattice=state.get("lattice"), q=state.get("q") if state.get("q") is not None else int(f.shape[0]),
            grid=tuple(grid) if grid is not None else None, tau=float(tau) if tau is not None else None,
            precision_policy=state.get("precision_policy"), storage_dtype=dtype, storage_offset=offset,
            collision=state.get("collision"), boundary=state.get("boundary"))

    def conflicts(self, solver_side: "RunIdentity", keys: tuple[str, ...]) -> list[str]:
        """Messages for each field in ``keys`` where this (saved) identity and the solver's disagree."""
        out = []
        for key in keys:
            saved, mine = getattr(self, key), getattr(solver_side, key)
            if key == "precision_policy":
                differ = saved != mine  # None is a tier of its own: plain dtype, no policy
            elif key == "storage_offset":
                differ = saved != mine
            elif saved is None or mine is None:
                continue
            elif key in ("lattice", "collision", "boundary"):
                differ = str(saved).upper() != str(mine).upper()
            else:
                differ = saved != mine
            if differ:
                out.append(_CONFLICT_TEXT[key].format(saved=saved, mine=mine))
        return out


_CONFLICT_TEXT = {
    "lattice": "the state's lattice {saved!r} is not the solver's lattice {mine!r}",
    "q": "the state has Q {saved} directions, the solver Q {mine}",
    "grid": "the state's grid {saved} is not the solver's grid {mine}",
    "tau": "the state's tau {saved} is not the solver's tau {mine}",
    "precision_policy": "the state's precision policy {saved!r} is not the solver's {mine!r}",
    "storage_dtype": "the state's storage dtype {saved} is not the solver's dtype {mine}",
    "storage_offset": ("the state's storage offset {saved!r} is not the solver's {mine!r} (deviation storage keeps "
                       "f - w); convert with absolute_populations() or build the solver with the matching "
                       "deviation flag"),
    "collision": "the state's collision {saved!r} is not the solver's collision {mine!r}",
    "boundary": "the state's boundary {saved!r} is not the solver's boundary {mine!r}",
}
_EAGER_KEYS = ("lattice", "q", "grid", "tau")
_FUSED_KEYS = ("precision_policy", "storage_dtype", "storage_offset", "collision", "boundary") + _EAGER_KEYS


# ------------------------------------------------------------------ the solver envelope


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


class _Rng:
    """torch RNG states in the envelope's ``rng`` entry (keys ``cpu``, ``cuda``, ``cuda_device``)."""

    @staticmethod
    def capture(device: torch.device) -> dict[str, Any]:
        snap: dict[str, Any] = {"cpu": torch.get_rng_state()}
        if device.type == "cuda":
            snap.update(cuda=torch.cuda.get_rng_state(device), cuda_device=str(device))
        return snap

    @staticmethod
    def apply(snap: object) -> None:
        if not (isinstance(snap, dict) and isinstance(snap.get("cpu"), torch.Tensor)):
            return  # nothing usable was recorded
        gpu_state, where = snap.get("cuda"), snap.get("cuda_device")
        if isinstance(gpu_state, torch.Tensor):
            idx = torch.device(where).index if isinstance(where, str) else None
            have = torch.cuda.is_available() and (idx is None or idx < torch.cuda.device_count())
            if not have:
                raise CheckpointError(f"the checkpoint's RNG state is for GPU {where or 'current'}, which this host "
                                      "lacks; pass restore_rng=False to skip it (a continuation that draws random "
                                      "numbers is then not bit-identical)")
        torch.set_rng_state(snap["cpu"])
        if isinstance(gpu_state, torch.Tensor):
            torch.cuda.set_rng_state(gpu_state, where)


def save_solver_checkpoint(path: str | Path, state: Mapping[str, Any], *,
                           metadata: Mapping[str, Any] | None = None) -> Path:
    """Write a state dict (from an adapter) as one tagged envelope with tensor digests; returns the path.

    Tensors are stored as dense CPU copies, so the file loads on any device (checkpoint.py:296-338).
    """
    body = _checked_state(state)
    device = body["populations"].device
    digests = {}
    for key in DIGESTED_KEYS:
        if isinstance(body.get(key), torch.Tensor):
            body[key] = _dense_cpu(body[key])
            digests[key] = _sha256(body[key])
    grid = body.get("grid")
    envelope = {
        "format": SOLVER_CHECKPOINT_FORMAT, "format_version": SOLVER_CHECKPOINT_VERSION,
        "created_at": datetime.now(timezone.utc).isoformat(), "step": int(body["step"]),
        "lattice": body.get("lattice"), "grid": None if grid is None else tuple(grid), "q": body.get("q"),
        "state": body, "metadata": dict(metadata or {}), "rng": _Rng.capture(device), "digests": digests,
    }
    return atomic_torch_save(envelope, path)


def _open_envelope(path: Path) -> dict:
    if not path.is_file():
        raise CheckpointError(f"checkpoint file not found: {path}")
    try:
        envelope = torch.load(path, map_location="cpu", weights_only=True)
    except Exception as exc:  # whatever the decoder trips on, the file is unusable
        raise CheckpointError(f"{path} could not be decoded (truncated or corrupt write?): "
                              f"{type(exc).__name__}: {exc}") from exc
    if not (isinstance(envelope, dict) and envelope.get("format") == SOLVER_CHECKPOINT_FORMAT):
        raise CheckpointError(f"{path} is not a solver checkpoint: missing format tag {SOLVER_CHECKPOINT_FORMAT!r}")
    if envelope.get("format_version") != SOLVER_CHECKPOINT_VERSION:
        raise CheckpointError(f"{path}: format_version={envelope.get('format_version')!r} is not supported "
                              f"(this reader knows {SOLVER_CHECKPOINT_VERSION})")
    return envelope


def _verify_digests(path: Path, state: Mapping[str, Any], record: object) -> None:
    if not isinstance(record, dict):
        return
    for key, want in record.items():
        value = state.get(key)
        if not (isinstance(value, torch.Tensor) and isinstance(want, str) and len(want) == 64):
            raise CheckpointError(f"{path} has an invalid integrity record for {key!r}")
        if _sha256(value) != want:
            raise CheckpointError(f"{path} failed the integrity digest for {key!r}: the stored tensor is not "
                                  "the one that was saved")


def load_solver_checkpoint(path: str | Path, *, expected_lattice: str | None = None,
                           expected_grid: tuple[int, ...] | None = None, expected_q: int | None = None,
                           restore_rng: bool = True) -> SolverCheckpoint:
    """Read and verify an envelope written by :func:`save_solver_checkpoint` (checkpoint.py:341-421).

    CheckpointError for a missing file, one that does not decode, has no format
    tag or another version, fails a digest, or differs from an ``expected_*``
    identity.  ``restore_rng`` puts the saved torch RNG states back.
    """
    target = Path(path)
    envelope = _open_envelope(target)
    state = _checked_state(envelope.get("state"))
    _verify_digests(target, state, envelope.get("digests"))
    found = RunIdentity.of_state(state)
    wanted = RunIdentity(lattice=expected_lattice, q=expected_q,
                         grid=tuple(expected_grid) if expected_grid is not None else None)
    for key, label in (("lattice", "lattice"), ("grid", "grid"), ("q", "Q")):
        want = getattr(wanted, key)
        got = found.lattice if key == "lattice" else (found.grid if key == "grid" else state.get("q"))
        if want is not None and got != want:
            raise CheckpointError(f"checkpoint {label} mismatch: expected {want!r}, found {got!r}")
    if restore_rng:
        _Rng.apply(envelope.get("rng"))
    q = state.get("q")
    return SolverCheckpoint(state=state, step=int(state["step"]),
                            lattice=found.lattice if isinstance(found.lattice, str) else None, grid=found.grid,
                            q=q if _plain_int(q) else None, metadata=dict(envelope.get("metadata") or {}),
                            format_version=SOLVER_CHECKPOINT_VERSION, path=target)


# ------------------------------------------------------------------ adapters


def _obstacle_of(solver: Any) -> torch.Tensor | None:
    for name in ("mask", "solid", "solid_mask"):
        v = getattr(solver, name, None)
        if isinstance(v, torch.Tensor):
            return v
    return None


def _state_of(solver: Any, f: torch.Tensor, step: int | None, extra, *, fused: bool) -> dict[str, Any]:
    if not isinstance(f, torch.Tensor):
        raise TypeError(f"populations must be a torch.Tensor, got {type(f).__name__}")
    ident = RunIdentity.of_solver(solver, f)
    counter = getattr(solver, "_report_step", 0)
    state: dict[str, Any] = {
        "populations": f.detach(),
        "step": int(step) if step is not None else int(counter or 0),
        "lattice": ident.lattice,
        "grid": ident.grid,
        "q": ident.q if ident.q is not None else int(f.shape[0]),
        "tau": ident.tau,
        "dtype": str(f.dtype),
        "obstacle_mask": _obstacle_of(solver),
        "precision_policy": ident.precision_policy,
    }
    if extra:
        state["extra"] = dict(extra)
    if fused:
        state["lattice"] = state["lattice"] or "D3Q19"
        state["storage_dtype"] = _encode_storage(ident.storage_dtype or f.dtype, ident.storage_offset)
        state["storage_offset"] = ident.storage_offset
        for key in ("collision", "boundary"):
            if getattr(ident, key) is not None:
                state[key] = getattr(ident, key)
    return state


def _restored(solver: Any, state: Mapping[str, Any], device, keys: tuple[str, ...]) -> torch.Tensor:
    s = _checked_state(state)
    f = s["populations"]
    saved, mine = RunIdentity.of_state(s), RunIdentity.of_solver(solver)
    problems = saved.conflicts(mine, keys)
    # the tensor itself must agree with the record and with the solver
    if not problems and mine.q is not None and int(f.shape[0]) != mine.q:
        problems.append(_CONFLICT_TEXT["q"].format(saved=int(f.shape[0]), mine=mine.q))
    if not problems and mine.grid is not None and tuple(f.shape[1:]) != mine.grid:
        problems.append(_CONFLICT_TEXT["grid"].format(saved=tuple(f.shape[1:]), mine=mine.grid))
    if not problems and "storage_dtype" in keys and mine.storage_dtype is not None and str(f.dtype) != mine.storage_dtype:
        problems.append(f"the state's populations are {f.dtype} but the record says {saved.storage_dtype}")
    if problems:
        raise CheckpointError("; ".join(problems))
    if _plain_int(getattr(solver, "_report_step", None)):
        solver._report_step = int(s["step"])
    target = device if device is not None else getattr(solver, "device", None)
    return f.to(device=target) if target is not None else f


def eager_state_dict(solver: Any, f: torch.Tensor, *, step: int | None = None,
                     extra: Mapping[str, Any] | None = None) -> dict[str, Any]:
    """State dict of an eager solver (``LBMStepExecutor`` or anything with ``tau`` and grid attributes).

    ``step`` defaults to the solver's ``run()`` counter (checkpoint.py:520-547).
    """
    return _state_of(solver, f, step, extra, fused=False)


def eager_load_state_dict(solver: Any, state: Mapping[str, Any], *,
                          device: torch.device | str | None = None) -> torch.Tensor:
    """Check lattice / Q / grid / τ, reset the solver's step counter, return the populations (checkpoint.py:550-565)."""
    return _restored(solver, state, device, _EAGER_KEYS)


def fused_state_dict(solver: Any, f: torch.Tensor, *, step: int | None = None,
                     extra: Mapping[str, Any] | None = None) -> dict[str, Any]:
    """State dict of a :class:`~tensorlbm_amd.FusedSolver3D` (``triton_fused_state_dict``, checkpoint.py:568-586).

    Adds the stored format (``storage_dtype``, ``storage_offset``) and, where the
    solver names them, its collision family and boundary.
    """
    return _state_of(solver, f, step, extra, fused=True)


def fused_load_state_dict(solver: Any, state: Mapping[str, Any], *,
                          device: torch.device | str | None = None) -> torch.Tensor:
    """Restore onto a :class:`~tensorlbm_amd.FusedSolver3D`, fail-closed (checkpoint.py:589-617).

    Checks the precision tier, the stored format (of the record and of the
    tensor), the collision family and boundary where both sides name them, and
    lattice / Q / grid / τ.
    """
    return _restored(solver, state, device, _FUSED_KEYS)


# the reference's adapter names (checkpoint.py:568, 589)
triton_fused_state_dict = fused_state_dict
triton_fused_load_state_dict = fused_load_state_dict


def absolute_populations(state: Mapping[str, Any]) -> torch.Tensor:
    """The state's populations as plain ``f_i`` in fp32 (``w_i`` added back for ``f − w`` storage).

    The fused kernels widen a stored deviation to fp32 and add the fp32 weight,
    so this is the value the next step starts from.
    """
    s = _checked_state(state)
    f = s["populations"].float()
    _, offset = _decode_storage(s.get("storage_dtype") or s.get("dtype"), s.get("storage_offset"))
    if offset == OFFSET_NONE:
        return f
    from .lattice import BY_Q

    q = int(f.shape[0])
    if q not in BY_Q:
        raise CheckpointError(f"no lattice with {q} directions for the {offset!r} offset")
    return f + BY_Q[q].W.to(f.device).reshape((q,) + (1,) * (f.ndim - 1))

