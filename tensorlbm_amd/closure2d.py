"""The reference's 2-D learned eddy-viscosity closure, on the GPU kernels of ``closure2d.hip``.

Same surface as the reference's ``ai`` package for D2Q9 LES (model.py, dataset.py, train.py,
inference.py):

* ``ModelArch`` / ``EddyViscosityMLP`` (model.py:19-92): in → H → … → H → 1, tanh / relu / gelu,
  Softplus(β = 10); ``set_feature_stats``; ``state_dict`` with the reference's keys
  (``feature_mean``, ``feature_std``, ``net.<i>.weight`` / ``bias``), so a model saved by either
  package loads in the other (``save_model`` / ``load_model``, model.py:99-142, tensor file plus
  JSON metadata, ``format_version`` 2).
* ``strain_rate_tensor_2d`` and ``extract_les_samples_2d`` (dataset.py:41-127), ``EddyViscosityDataset``
  with ``split``, ``save_dataset_pt`` / ``load_dataset_pt`` (dataset.py:153-229).
* ``train_eddy_viscosity_model`` (train.py:310-464): the same split, initialisation, feature
  statistics, shuffled minibatches, MSE, gradient-norm clipping, Adam, learning-rate schedules,
  patience and best-validation state; forward and backward passes are the HIP kernels.
* ``predict_nu_t_2d``, ``predict_tau_eff_2d`` and ``collide_ai_les_bgk`` (inference.py:26-92).

Computation runs on the GPU in float32; there is no CPU path.  The hidden width must be 8, 16,
32 or 64 and the depth 1-4 (training: width × depth ≤ 120).
"""

from __future__ import annotations

import ctypes
import json
import math
import time
from dataclasses import asdict, dataclass
from pathlib import Path

import torch

from . import _native as nat

__all__ = [
    "ModelArch", "EddyViscosityMLP", "save_model", "load_model", "strain_rate_tensor_2d", "extract_les_samples_2d",
    "extract_les_samples_2d_multi", "EddyViscosityDataset", "save_dataset_pt", "load_dataset_pt", "TrainConfig",
    "train_eddy_viscosity_model", "predict_nu_t_2d", "predict_tau_eff_2d", "collide_ai_les_bgk",
]

_ACTIVATIONS = {"tanh": 0, "relu": 1, "gelu": 2}


class _Arch(ctypes.Structure):
    _fields_ = [("in_features", ctypes.c_int32), ("hidden_features", ctypes.c_int32),
                ("hidden_layers", ctypes.c_int32), ("activation", ctypes.c_int32)]


@dataclass(frozen=True)
class ModelArch:
    """Hyper-parameters of an :class:`EddyViscosityMLP` (model.py:19-27)."""

    in_features: int = 3
    hidden_features: int = 16
    n_hidden_layers: int = 2
    activation: str = "tanh"

    def native(self) -> _Arch:
        name = self.activation.lower()
        if name not in _ACTIVATIONS:
            raise ValueError(f"Unsupported activation: {self.activation!r}")
        return _Arch(int(self.in_features), int(self.hidden_features), int(self.n_hidden_layers), _ACTIVATIONS[name])


def _gpu(device) -> torch.device:
    d = torch.device(device)
    if d.type != "cuda":
        raise RuntimeError(f"the learned closure runs on a GPU device (the HIP kernels), got {d}")
    return d


class EddyViscosityMLP:
    """MLP from strain features to a non-negative eddy viscosity, held as one fp32 device vector.

    The vector is the trainable parameters in ``nn.Sequential`` order (each hidden layer's weight
    and bias, then the output layer's), followed by the input mean and std; ``params`` is that
    vector.  A new model draws its initial weights as ``torch.nn.Linear`` does, from torch's
    global generator, layer by layer in order: after ``torch.manual_seed(s)`` it starts from the
    same weights as the reference's ``EddyViscosityMLP(arch)`` (model.py:50-73).
    """

    def __init__(self, arch: ModelArch | None = None, *, device: str | torch.device = "cuda:0") -> None:
        self.arch = arch or ModelArch()
        self._native = self.arch.native()
        self.device = _gpu(device)
        self.nparam = int(nat.lib().tlbm_mlp2d_nparam(ctypes.byref(self._native)))
        if self.nparam < 0:
            nat.check(nat.ERR_UNSUPPORTED, "EddyViscosityMLP")
        a = self.arch
        pieces, in_dim = [], a.in_features
        for _ in range(a.n_hidden_layers):
            lin = torch.nn.Linear(in_dim, a.hidden_features)
            pieces += [lin.weight.detach().reshape(-1), lin.bias.detach()]
            in_dim = a.hidden_features
        lin = torch.nn.Linear(in_dim, 1)
        pieces += [lin.weight.detach().reshape(-1), lin.bias.detach(), torch.zeros(a.in_features),
                   torch.ones(a.in_features)]
        self.params = torch.cat(pieces).to(device=self.device, dtype=torch.float32).contiguous()
        assert self.params.numel() == self.nparam + 2 * a.in_features
        self.training = True

    # -- torch-module-like surface --------------------------------------------
    def eval(self):
        self.training = False
        return self

    def train(self, mode: bool = True):
        self.training = bool(mode)
        return self

    def to(self, device=None, dtype=None):
        if dtype is not None and dtype != torch.float32:
            raise ValueError("EddyViscosityMLP computes in float32")
        if device is not None:
            self.params = self.params.to(_gpu(device))
            self.device = self.params.device
        return self

    @property
    def feature_mean(self) -> torch.Tensor:
        return self.params[self.nparam:self.nparam + self.arch.in_features]

    @property
    def feature_std(self) -> torch.Tensor:
        return self.params[self.nparam + self.arch.in_features:]

    def set_feature_stats(self, mean: torch.Tensor, std: torch.Tensor) -> None:
        """Persist feature-normalisation statistics (std floored at 1e-6), model.py:66-80."""
        n = self.arch.in_features
        mean, std = torch.as_tensor(mean), torch.as_tensor(std)
        if tuple(mean.shape) != (n,) or tuple(std.shape) != (n,):
            raise ValueError(f"Normalization stats must have shape ({n},), got {tuple(mean.shape)} "
                             f"and {tuple(std.shape)}")
        self.params[self.nparam:self.nparam + n] = mean.detach().to(self.device, torch.float32)
        self.params[self.nparam + n:] = std.detach().to(self.device, torch.float32).clamp_min(1e-6)

    def _layers(self):
        """(name, offset, shape) of every weight and bias in nn.Sequential numbering."""
        a, out, off, in_dim = self.arch, [], 0, self.arch.in_features
        for k in range(a.n_hidden_layers + 1):
            rows = a.hidden_features if k < a.n_hidden_layers else 1
            out.append((f"net.{2 * k}.weight", off, (rows, in_dim)))
            off += rows * in_dim
            out.append((f"net.{2 * k}.bias", off, (rows,)))
            off += rows
            in_dim = a.hidden_features
        return out

    def state_dict(self) -> dict:
        """The reference module's state dict (CPU tensors): normalisation buffers and layer tensors."""
        p = self.params.detach().cpu()
        sd = {"feature_mean": self.feature_mean.detach().cpu().clone(),
              "feature_std": self.feature_std.detach().cpu().clone()}
        for name, off, shape in self._layers():
            n = math.prod(shape)
            sd[name] = p[off:off + n].reshape(shape).clone()
        return sd

    def load_state_dict(self, sd: dict) -> None:
        want = {"feature_mean", "feature_std"} | {name for name, _, _ in self._layers()}
        if set(sd) != want:
            raise ValueError(f"state dict keys {sorted(sd)} do not match the architecture's {sorted(want)}")
        vec = torch.empty_like(self.params, device="cpu")
        for name, off, shape in self._layers():
            t = torch.as_tensor(sd[name])
            if tuple(t.shape) != shape:
                raise ValueError(f"{name}: shape {tuple(t.shape)} != {shape}")
            vec[off:off + math.prod(shape)] = t.reshape(-1).float()
        n = self.arch.in_features
        vec[self.nparam:self.nparam + n] = torch.as_tensor(sd["feature_mean"]).float()
        vec[self.nparam + n:] = torch.as_tensor(sd["feature_std"]).float()
        self.params.copy_(vec.to(self.device))

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return self.forward(x)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """ν_t of features ``x`` (N, in_features) → (N, 1), model.py:83-92."""
        if x.shape[-1] != self.arch.in_features:
            raise ValueError(f"Expected input with {self.arch.in_features} features, got tensor of shape "
                             f"{tuple(x.shape)}")
        nat.require_gpu(x, "EddyViscosityMLP")
        lead = tuple(x.shape[:-1])
        xf = x.to(device=self.device, dtype=torch.float32).reshape(-1, self.arch.in_features).contiguous()
        y = torch.empty((xf.shape[0], 1), dtype=torch.float32, device=self.device)
        nat.check(nat.lib().tlbm_mlp2d_forward(ctypes.byref(self._native), self.params.data_ptr(), xf.data_ptr(),
                                               int(xf.shape[0]), y.data_ptr(), nat.stream_of(xf)),
                  "EddyViscosityMLP")
        return y.reshape(*lead, 1)


def save_model(model: EddyViscosityMLP, path) -> Path:
    """Tensor-only state dict plus ``<path>.json`` metadata, as model.py:99-118 writes them."""
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    torch.save(model.state_dict(), p)
    meta = {"arch": asdict(model.arch),
            "normalization": {"feature_mean": model.feature_mean.detach().cpu().tolist(),
                              "feature_std": model.feature_std.detach().cpu().tolist()},
            "format_version": 2}
    p.with_suffix(p.suffix + ".json").write_text(json.dumps(meta, indent=2))
    return p


def load_model(path, device: str | torch.device = "cuda:0") -> EddyViscosityMLP:
    """Inverse of :func:`save_model` for files of either package (model.py:121-142), on ``device``."""
    p = Path(path)
    blob = torch.load(p, map_location="cpu", weights_only=True)
    meta_path = p.with_suffix(p.suffix + ".json")
    if meta_path.exists():
        arch_dict = json.loads(meta_path.read_text()).get("arch") or {}
    elif isinstance(blob, dict) and "arch" in blob:
        arch_dict = blob.get("arch") or {}
    else:
        arch_dict = {}
    model = EddyViscosityMLP(ModelArch(**arch_dict) if arch_dict else ModelArch(), device=device)
    sd = blob["state_dict"] if isinstance(blob, dict) and "state_dict" in blob else blob
    if not isinstance(sd, dict):
        raise ValueError(f"Unsupported model payload in {p}")
    model.load_state_dict(sd)
    return model.eval()


# ---------------------------------------------------------------------------
# features and samples
# ---------------------------------------------------------------------------


def _velocity_pair(ux, uy, what):
    if ux.shape != uy.shape or ux.ndim != 2:
        raise ValueError(f"ux and uy must be 2-D tensors of equal shape, got {tuple(ux.shape)} and "
                         f"{tuple(uy.shape)}")
    nat.require_gpu(ux, what)
    return ux.float().contiguous(), uy.to(ux.device, torch.float32).contiguous()


def _strain(ux, uy, c_s, what, with_target):
    ux, uy = _velocity_pair(ux, uy, what)
    ny, nx = ux.shape
    feats = torch.empty((ny, nx, 3), dtype=torch.float32, device=ux.device)
    target = torch.empty((ny, nx), dtype=torch.float32, device=ux.device) if with_target else None
    nat.check(nat.lib().tlbm_strain2d(ux.data_ptr(), uy.data_ptr(), ny, nx, feats.data_ptr(), nat.ptr(target),
                                      float(c_s), nat.stream_of(ux)), what)
    return feats, target


def strain_rate_tensor_2d(ux: torch.Tensor, uy: torch.Tensor):
    """(S_xx, S_yy, S_xy) of a periodic 2-D velocity field by central differences (dataset.py:41-72)."""
    feats, _ = _strain(ux, uy, 0.0, "strain_rate_tensor_2d", False)
    return feats[..., 0], feats[..., 1], feats[..., 2]


def extract_les_samples_2d(ux: torch.Tensor, uy: torch.Tensor, c_s: float = 0.1, mask: torch.Tensor | None = None):
    """(features (N, 3), target (N, 1)) with target = c_s² |S| (Δ = 1), solid cells of ``mask`` removed
    (dataset.py:89-127)."""
    feats, target = _strain(ux, uy, c_s, "extract_les_samples_2d", True)
    f, t = feats.reshape(-1, 3), target.reshape(-1, 1)
    if mask is not None:
        if mask.shape != ux.shape:
            raise ValueError(f"mask shape {tuple(mask.shape)} must match velocity shape {tuple(ux.shape)}")
        keep = (~mask.to(f.device, torch.bool).reshape(-1)).nonzero(as_tuple=False).squeeze(-1)
        f, t = f.index_select(0, keep), t.index_select(0, keep)
    return f.contiguous(), t.contiguous()


def extract_les_samples_2d_multi(snapshots, c_s: float = 0.1, mask: torch.Tensor | None = None):
    """Samples of many ``(ux, uy)`` snapshots concatenated; empty (0, 3) / (0, 1) tensors when there are
    none (dataset.py:130-144)."""
    rows = [extract_les_samples_2d(ux, uy, c_s=c_s, mask=mask) for ux, uy in snapshots]
    if not rows:
        return torch.zeros(0, 3), torch.zeros(0, 1)
    return torch.cat([r[0] for r in rows], dim=0), torch.cat([r[1] for r in rows], dim=0)


@dataclass
class EddyViscosityDataset:
    """In-memory (features, target) samples plus metadata (dataset.py:153-191)."""

    features: torch.Tensor
    targets: torch.Tensor
    c_s: float = 0.1
    description: str = ""

    def __len__(self) -> int:
        return int(self.features.shape[0])

    def split(self, val_fraction: float = 0.1, seed: int = 0):
        """Seeded random (train, val) split, ``max(1, round(n · val_fraction))`` validation samples."""
        if not 0.0 < val_fraction < 1.0:
            raise ValueError("val_fraction must be in (0, 1)")
        n = len(self)
        n_val = max(1, int(round(n * val_fraction)))
        perm = torch.randperm(n, generator=torch.Generator().manual_seed(int(seed)))
        parts = []
        for idx, tag in ((perm[n_val:], " [train]"), (perm[:n_val], " [val]")):
            i = idx.to(self.features.device)
            parts.append(EddyViscosityDataset(self.features.index_select(0, i), self.targets.index_select(0, i),
                                              self.c_s, self.description + tag))
        return parts[0], parts[1]


def save_dataset_pt(dataset: EddyViscosityDataset, path) -> Path:
    """Tensor file of features and targets plus ``<path>.json`` metadata (dataset.py:194-205)."""
    p = Path(path)
    p.parent.mkdir(parents=True, exist_ok=True)
    torch.save({"features": dataset.features.detach().cpu(), "targets": dataset.targets.detach().cpu()}, p)
    p.with_suffix(p.suffix + ".json").write_text(json.dumps(
        {"c_s": float(dataset.c_s), "description": dataset.description, "format_version": 2}, indent=2))
    return p


def load_dataset_pt(path) -> EddyViscosityDataset:
    """Inverse of :func:`save_dataset_pt` (dataset.py:208-229); tensors load on the CPU."""
    p = Path(path)
    blob = torch.load(p, map_location="cpu", weights_only=True)
    meta_path = p.with_suffix(p.suffix + ".json")
    if meta_path.exists():
        meta = json.loads(meta_path.read_text())
    elif isinstance(blob, dict):
        meta = {"c_s": blob.get("c_s", 0.1), "description": blob.get("description", "")}
    else:
        meta = {}
    if not isinstance(blob, dict) or "features" not in blob or "targets" not in blob:
        raise ValueError(f"Unsupported dataset payload in {p}")
    return EddyViscosityDataset(blob["features"], blob["targets"], float(meta.get("c_s", 0.1)),
                                str(meta.get("description", "")))


# ---------------------------------------------------------------------------
# training
# ---------------------------------------------------------------------------


@dataclass
class TrainConfig:
    """Hyper-parameters of :func:`train_eddy_viscosity_model` (train.py:31-46); ``device`` is a GPU."""

    epochs: int = 20
    batch_size: int = 4096
    learning_rate: float = 1e-3
    val_fraction: float = 0.1
    seed: int = 0
    hidden_features: int = 16
    n_hidden_layers: int = 2
    activation: str = "tanh"
    device: str = "cuda:0"
    lr_scheduler: str = "none"
    patience: int | None = None
    gradient_clip_norm: float | None = 1.0


class _Plateau:
    """ReduceLROnPlateau(mode="min", factor=0.5, patience=2) with torch's defaults (relative
    threshold 1e-4, no cooldown, min_lr 0, eps 1e-8)."""

    def __init__(self, lr: float) -> None:
        self.lr, self.best, self.bad = lr, math.inf, 0

    def step(self, metric: float) -> None:
        if metric < self.best * (1.0 - 1e-4):
            self.best, self.bad = metric, 0
        else:
            self.bad += 1
        if self.bad > 2:
            new = self.lr * 0.5
            if self.lr - new > 1e-8:
                self.lr = new
            self.bad = 0


def _lr_schedule(cfg: TrainConfig):
    name = str(cfg.lr_scheduler).strip().lower()
    if name not in ("none", "cosine", "plateau"):
        raise ValueError(f"Unsupported lr_scheduler: {cfg.lr_scheduler!r}")
    return name


def train_eddy_viscosity_model(dataset, out_path, config: TrainConfig | None = None, *,
                               backend: str | None = None) -> dict:
    """Train an :class:`EddyViscosityMLP` on the GPU kernels and save it to ``out_path`` (train.py:310-464).

    Returns the reference's metadata dict (history, final metrics of the best epoch, arch, config,
    path) with ``backend`` "hip".
    """
    if backend not in (None, "torch", "hip"):
        raise NotImplementedError(f"backend {backend!r} is not on the HIP path")
    cfg = config or TrainConfig()
    sched = _lr_schedule(cfg)
    if not isinstance(dataset, EddyViscosityDataset):
        dataset = load_dataset_pt(dataset)
    if len(dataset) < 4:
        raise ValueError(f"Dataset is too small to train ({len(dataset)} samples).")
    device = _gpu(cfg.device)
    train_ds, val_ds = dataset.split(cfg.val_fraction, seed=cfg.seed)
    x_train = train_ds.features.to(device, torch.float32).contiguous()
    y_train = train_ds.targets.to(device, torch.float32).reshape(-1).contiguous()
    x_val = val_ds.features.to(device, torch.float32).contiguous()
    y_val = val_ds.targets.to(device, torch.float32).reshape(-1, 1)

    torch.manual_seed(int(cfg.seed))
    arch = ModelArch(int(x_train.shape[-1]), int(cfg.hidden_features), int(cfg.n_hidden_layers), str(cfg.activation))
    model = EddyViscosityMLP(arch, device=device)
    model.set_feature_stats(x_train.mean(dim=0), x_train.std(dim=0, unbiased=False).clamp_min(1e-6))

    lib, s, na = nat.lib(), nat.stream_of(x_train), ctypes.byref(model._native)
    n_train = int(x_train.shape[0])
    batch_size = max(1, min(int(cfg.batch_size), n_train))
    nblk_max = (batch_size + 63) // 64
    partials = torch.empty((nblk_max, model.nparam + 2), dtype=torch.float32, device=device)
    m = torch.zeros(model.nparam, dtype=torch.float32, device=device)
    v = torch.zeros_like(m)
    stats = torch.zeros(2, dtype=torch.float32, device=device)
    clip = float(cfg.gradient_clip_norm) if cfg.gradient_clip_norm is not None else 0.0
    lr0, lr = float(cfg.learning_rate), float(cfg.learning_rate)
    plateau = _Plateau(lr0) if sched == "plateau" else None
    generator = torch.Generator(device="cpu").manual_seed(int(cfg.seed))
    history: list[dict] = []
    best_params, best_epoch, best_val, since_best = model.params.clone(), 0, math.inf, 0
    step = 0
    t0 = time.perf_counter()
    for epoch in range(int(cfg.epochs)):
        sums = torch.zeros(2, dtype=torch.float64, device=device)
        perm = torch.randperm(n_train, generator=generator)
        for i in range(0, n_train, batch_size):
            idx = perm[i:i + batch_size].to(device)
            xb = x_train.index_select(0, idx).contiguous()
            yb = y_train.index_select(0, idx).contiguous()
            nb = int(xb.shape[0])
            step += 1
            nat.check(lib.tlbm_mlp2d_grad(na, model.params.data_ptr(), xb.data_ptr(), yb.data_ptr(), nb,
                                          partials.data_ptr(), s), "train_eddy_viscosity_model")
            nat.check(lib.tlbm_mlp2d_adam(na, model.params.data_ptr(), partials.data_ptr(), (nb + 63) // 64,
                                          m.data_ptr(), v.data_ptr(), step, lr, 0.9, 0.999, 1e-8, clip,
                                          stats.data_ptr(), s), "train_eddy_viscosity_model")
            sums[0] += stats[0].double() * nb
            sums[1] += stats[1].double()
        train_mse, train_mae = (sums / max(1, n_train)).tolist()
        val_pred = model(x_val)
        err = val_pred - y_val
        val_mse = float((err * err).mean())
        val_mae = float(err.abs().mean())
        var = float(torch.var(y_val, unbiased=False))
        val_r2 = 0.0 if var <= 0.0 else 1.0 - val_mse / var
        history.append({"epoch": epoch, "train_mse": train_mse, "train_mae": train_mae, "val_mse": val_mse,
                        "val_mae": val_mae, "val_r2": val_r2, "lr": lr})
        if val_mse < best_val - 1e-12:
            best_val, best_epoch, since_best = val_mse, epoch, 0
            best_params = model.params.clone()
        else:
            since_best += 1
        if sched == "cosine":
            lr = 0.5 * lr0 * (1.0 + math.cos(math.pi * (epoch + 1) / max(1, int(cfg.epochs))))
        elif plateau is not None:
            plateau.step(val_mse)
            lr = plateau.lr
        if cfg.patience is not None and since_best > int(cfg.patience):
            break
    model.params.copy_(best_params)
    model.eval()
    path = save_model(model, out_path)
    final = history[best_epoch] if history else {}
    nan = float("nan")
    return {
        "path": str(path), "arch": asdict(arch), "config": asdict(cfg), "backend": "hip",
        "n_samples_train": n_train, "n_samples_val": int(x_val.shape[0]), "history": history,
        "final_train_mse": float(final.get("train_mse", nan)), "final_train_mae": float(final.get("train_mae", nan)),
        "final_val_mse": float(final.get("val_mse", nan)), "final_val_mae": float(final.get("val_mae", nan)),
        "final_val_r2": float(final.get("val_r2", nan)), "best_epoch": int(best_epoch),
        "stopped_early": len(history) < int(cfg.epochs), "training_time_s": time.perf_counter() - t0,
    }


# ---------------------------------------------------------------------------
# inference
# ---------------------------------------------------------------------------


def _nu_t(model: EddyViscosityMLP, ux, uy, mode, tau, what):
    ux, uy = _velocity_pair(ux, uy, what)
    ny, nx = ux.shape
    out = torch.empty((ny, nx), dtype=torch.float32, device=ux.device)
    nat.check(nat.lib().tlbm_mlp2d_nu_t(ctypes.byref(model._native), model.params.data_ptr(), ux.data_ptr(),
                                        uy.data_ptr(), ny, nx, mode, float(tau), out.data_ptr(), nat.stream_of(ux)),
              what)
    return out


def predict_nu_t_2d(model: EddyViscosityMLP, ux: torch.Tensor, uy: torch.Tensor) -> torch.Tensor:
    """Non-negative ν_t field (ny, nx) of a velocity snapshot (inference.py:26-50)."""
    return _nu_t(model, ux, uy, 0, 0.0, "predict_nu_t_2d")


def predict_tau_eff_2d(model: EddyViscosityMLP, ux: torch.Tensor, uy: torch.Tensor, tau: float) -> torch.Tensor:
    """``max(τ + 3 ν_t, 0.5 + 1e-3)`` per cell (inference.py:53-65)."""
    return _nu_t(model, ux, uy, 1, tau, "predict_tau_eff_2d")


def collide_ai_les_bgk(f: torch.Tensor, tau: float, model: EddyViscosityMLP) -> torch.Tensor:
    """D2Q9 BGK with the learned τ_eff per cell (inference.py:68-92): macroscopic moments, then one
    kernel that takes each cell's strain, evaluates the network and relaxes."""
    what = "collide_ai_les_bgk"
    nat.require_gpu(f, what)
    if f.ndim != 3 or f.shape[0] != 9:
        raise ValueError(f"{what}: populations must have shape (9, ny, nx), got {tuple(f.shape)}")
    if f.dtype != torch.float32:
        raise ValueError(f"Unsupported dtype {f.dtype} for {what}; use float32")
    f = f.contiguous()
    ny, nx = int(f.shape[1]), int(f.shape[2])
    rho, ux, uy = (torch.empty((ny, nx), dtype=torch.float32, device=f.device) for _ in range(3))
    lib, s = nat.lib(), nat.stream_of(f)
    nat.check(lib.tlbm_macroscopic(9, nat.STORE_F32, f.data_ptr(), rho.data_ptr(), ux.data_ptr(), uy.data_ptr(),
                                   None, ny * nx, s), what)
    out = torch.empty_like(f)
    nat.check(lib.tlbm_ai_les_bgk(ctypes.byref(model._native), model.params.data_ptr(), f.data_ptr(), out.data_ptr(),
                                  rho.data_ptr(), ux.data_ptr(), uy.data_ptr(), ny, nx, float(tau), s), what)
    return out
