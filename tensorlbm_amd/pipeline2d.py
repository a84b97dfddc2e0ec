"""The reference's 2-D learned-closure pipeline on the HIP kernels: data run -> samples -> training ->
an LES run with the trained closure (``run_ai_les_pipeline``, ai/pipeline.py:222-396).

Every lattice operation is a library op on the GPU: D2Q9 equilibrium, BGK / Smagorinsky BGK collision,
periodic stream and moments (``tensorlbm_amd.ops``); strain, samples, training, ν_t and the AI-LES
collision (``tensorlbm_amd.closure2d``).  The initial field and the block average are a few torch
element-wise ops on the same device.

The reference also records each run, dataset and model in an SQLite run database (ai/database.py);
that bookkeeping is not part of this package, so ``AIPipelineResult.db_path`` and the three ids are
``None`` here and everything else (dataset and model files, training history, validation
diagnostics) is as the reference returns it.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any, Literal

import torch

from . import _native as nat
from . import ops
from .closure2d import (
    EddyViscosityDataset,
    TrainConfig,
    collide_ai_les_bgk,
    extract_les_samples_2d,
    load_model,
    predict_nu_t_2d,
    save_dataset_pt,
    train_eddy_viscosity_model,
)

__all__ = ["AIPipelineResult", "run_ai_les_pipeline", "run_ai_dns_pipeline"]


def _device(device) -> torch.device:
    d = torch.device(device)
    if d.type != "cuda":
        raise RuntimeError(f"the pipeline runs on a GPU device (the HIP kernels), got {d}; "
                           "the HIP path has no CPU fallback")
    nat.lib()  # fail here, before any work, when the library is missing
    return d


def _init_random_velocity_field(nx: int, ny: int, seed: int, device, mean_u: float = 0.05) -> torch.Tensor:
    """D2Q9 equilibrium at ρ = 1 of a superposition of low-wavenumber sinusoids (pipeline.py:35-65)."""
    torch.manual_seed(int(seed))  # the reference seeds torch's generator here; kept for the caller's RNG state
    yy, xx = torch.meshgrid(torch.arange(ny, device=device).float(), torch.arange(nx, device=device).float(),
                            indexing="ij")
    kx, ky = 2.0 * math.pi / max(nx, 1), 2.0 * math.pi / max(ny, 1)
    ux = mean_u + 0.02 * (torch.sin(2.0 * kx * xx) * torch.cos(ky * yy)
                          + 0.5 * torch.sin(4.0 * kx * xx + 0.3) * torch.cos(2.0 * ky * yy))
    uy = 0.02 * (torch.cos(kx * xx) * torch.sin(2.0 * ky * yy)
                 + 0.5 * torch.cos(3.0 * kx * xx) * torch.sin(ky * yy + 0.7))
    return ops.equilibrium(torch.ones_like(ux), ux, uy)


def _velocity(f: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    _, ux, uy = ops.macroscopic(f)
    return ux.clone(), uy.clone()


def _run_les_smoke(nx, ny, tau, c_s, n_steps, sample_every, seed, device):
    """Smagorinsky-BGK data run, ``(ux, uy)`` every ``sample_every`` steps (pipeline.py:68-90)."""
    f = _init_random_velocity_field(nx, ny, seed, device)
    snapshots = []
    for step in range(int(n_steps)):
        f = ops.stream(ops.collide_smagorinsky_bgk(f, float(tau), float(c_s)))
        if sample_every > 0 and (step + 1) % sample_every == 0:
            snapshots.append(_velocity(f))
    return snapshots or [_velocity(f)]


def _coarsen_mean(field: torch.Tensor, factor: int) -> torch.Tensor:
    """Block average of a 2-D field by an integer factor, trailing rows / columns dropped (pipeline.py:93-107)."""
    if factor <= 1:
        return field
    if field.ndim != 2:
        raise ValueError(f"field must be 2-D, got shape {tuple(field.shape)}")
    cy, cx = int(field.shape[0]) // int(factor), int(field.shape[1]) // int(factor)
    if cy < 1 or cx < 1:
        raise ValueError(f"coarsen factor {factor} is too large for field shape {tuple(field.shape)}")
    return field[:cy * factor, :cx * factor].reshape(cy, factor, cx, factor).mean(dim=(1, 3))


def _run_dns_reference(nx, ny, tau, n_steps, sample_every, seed, device, dns_scale=2, warmup_steps=20):
    """Plain BGK on a ``dns_scale``-times finer grid, block-averaged snapshots (pipeline.py:110-149)."""
    scale = max(1, int(dns_scale))
    f = _init_random_velocity_field(int(nx) * scale, int(ny) * scale, seed, device)
    for _ in range(max(0, int(warmup_steps))):
        f = ops.stream(ops.collide_bgk(f, float(tau)))

    def coarse():
        ux, uy = _velocity(f)
        return _coarsen_mean(ux, scale), _coarsen_mean(uy, scale)

    snapshots = []
    for step in range(int(n_steps)):
        f = ops.stream(ops.collide_bgk(f, float(tau)))
        if sample_every > 0 and (step + 1) % sample_every == 0:
            snapshots.append(coarse())
    return snapshots or [coarse()]


def _run_ai_validation(nx, ny, tau, n_steps, seed, device, model_path) -> dict[str, Any]:
    """AI-LES and plain BGK side by side from the same field; max |u|, kinetic energy, ν_t statistics and
    finiteness of the AI-LES run (pipeline.py:152-187)."""
    model = load_model(model_path, device=device)
    f_ai = _init_random_velocity_field(nx, ny, seed, device)
    f_bgk = f_ai.clone()
    for _ in range(int(n_steps)):
        f_ai = ops.stream(collide_ai_les_bgk(f_ai, float(tau), model))
        f_bgk = ops.stream(ops.collide_bgk(f_bgk, float(tau)))
    ux_ai, uy_ai = _velocity(f_ai)
    ux_bgk, uy_bgk = _velocity(f_bgk)
    nu_t = predict_nu_t_2d(model, ux_ai, uy_ai)
    sq_ai, sq_bgk = ux_ai * ux_ai + uy_ai * uy_ai, ux_bgk * ux_bgk + uy_bgk * uy_bgk
    return {
        "ai_umax": float(torch.sqrt(sq_ai).max()),
        "bgk_umax": float(torch.sqrt(sq_bgk).max()),
        "ai_ke": float((0.5 * sq_ai).mean()),
        "bgk_ke": float((0.5 * sq_bgk).mean()),
        "ai_nu_t_mean": float(nu_t.mean()),
        "ai_nu_t_max": float(nu_t.max()),
        "stable": bool(torch.isfinite(ux_ai).all() and torch.isfinite(uy_ai).all()),
    }


@dataclass
class AIPipelineResult:
    """Artefacts and diagnostics of :func:`run_ai_les_pipeline` (pipeline.py:195-219); ``db_path`` and
    the ids stay ``None`` (no run database in this package)."""

    work_dir: Path
    db_path: Path | None
    dataset_path: Path
    model_path: Path
    run_id: int | None
    dataset_id: int | None
    model_id: int | None
    n_samples: int
    data_source: str = "les"
    n_snapshots: int = 0
    training_time_s: float = 0.0
    training: dict[str, Any] = field(default_factory=dict)
    validation: dict[str, Any] = field(default_factory=dict)

    def to_dict(self) -> dict[str, Any]:
        d = self.__dict__.copy()
        for k in ("work_dir", "db_path", "dataset_path", "model_path"):
            d[k] = None if d[k] is None else str(d[k])
        return d


def run_ai_les_pipeline(work_dir, nx: int = 64, ny: int = 64, tau: float = 0.8, c_s: float = 0.1,
                        data_steps: int = 40, sample_every: int = 10, val_steps: int = 20,
                        train_config: TrainConfig | None = None, seed: int = 0, device: str = "cuda:0",
                        run_name: str = "ai_les_demo", data_source: Literal["les", "dns"] = "les",
                        dns_scale: int = 2, dns_warmup_steps: int = 20) -> AIPipelineResult:
    """Data run (Smagorinsky LES, or a finer BGK run block-averaged for ``data_source="dns"``) ->
    Smagorinsky-labelled samples -> ``dataset.pt`` -> training -> ``model.pt`` -> validation run with the
    trained closure from seed + 1 (pipeline.py:222-396).  ``run_name`` names the dataset's description."""
    work = Path(work_dir)
    work.mkdir(parents=True, exist_ok=True)
    source = str(data_source).strip().lower()
    if source not in {"les", "dns"}:
        raise ValueError(f"data_source must be 'les' or 'dns', got {data_source!r}")
    dev = _device(device)
    dns_scale, dns_warmup_steps = max(1, int(dns_scale)), max(0, int(dns_warmup_steps))
    if source == "dns":
        snapshots = _run_dns_reference(int(nx), int(ny), float(tau), int(data_steps), int(sample_every), int(seed),
                                       dev, dns_scale, dns_warmup_steps)
    else:
        snapshots = _run_les_smoke(int(nx), int(ny), float(tau), float(c_s), int(data_steps), int(sample_every),
                                   int(seed), dev)
    rows = [extract_les_samples_2d(ux, uy, c_s=float(c_s)) for ux, uy in snapshots]
    dataset = EddyViscosityDataset(torch.cat([r[0] for r in rows]), torch.cat([r[1] for r in rows]), float(c_s),
                                   f"{source.upper()} snapshots of {run_name} ({len(snapshots)} frames)")
    dataset_path = save_dataset_pt(dataset, work / "dataset.pt")
    model_path = work / "model.pt"
    cfg = train_config or TrainConfig()
    if torch.device(cfg.device) != dev:
        cfg = TrainConfig(**{**cfg.__dict__, "device": str(dev)})
    training = train_eddy_viscosity_model(dataset, model_path, cfg)
    validation = _run_ai_validation(int(nx), int(ny), float(tau), int(val_steps), int(seed) + 1, dev, model_path)
    return AIPipelineResult(work_dir=work, db_path=None, dataset_path=dataset_path, model_path=model_path,
                            run_id=None, dataset_id=None, model_id=None, n_samples=len(dataset),
                            data_source=source, n_snapshots=len(snapshots),
                            training_time_s=float(training.get("training_time_s", 0.0)), training=training,
                            validation=validation)


def run_ai_dns_pipeline(work_dir, nx: int = 64, ny: int = 64, tau: float = 0.8, c_s: float = 0.1,
                        data_steps: int = 40, sample_every: int = 10, val_steps: int = 20,
                        train_config: TrainConfig | None = None, seed: int = 0, device: str = "cuda:0",
                        run_name: str = "ai_dns_demo", dns_scale: int = 2,
                        dns_warmup_steps: int = 20) -> AIPipelineResult:
    """:func:`run_ai_les_pipeline` with ``data_source="dns"`` (pipeline.py:399-432)."""
    return run_ai_les_pipeline(work_dir, nx=nx, ny=ny, tau=tau, c_s=c_s, data_steps=data_steps,
                               sample_every=sample_every, val_steps=val_steps, train_config=train_config, seed=seed,
                               device=device, run_name=run_name, data_source="dns", dns_scale=dns_scale,
                               dns_warmup_steps=dns_warmup_steps)
