"""Learned eddy-viscosity LES closure on the GPU (3-D counterpart of ``tensorlbm.ai``).

The reference's closure family is 2-D: ``EddyViscosityMLP`` (src/tensorlbm/ai/model.py:19-92)
maps strain features to ν_t ≥ 0 through ``Linear`` layers and ``Softplus(β=10)``,
``train_eddy_viscosity_model`` (ai/train.py:310-464) fits it with MSE + Adam and gradient-norm
clipping, and ``collide_ai_les_bgk`` (ai/inference.py:68-92) relaxes with
``τ_eff = max(τ + 3 ν_t, 0.501)``.  Here the network is the 3-D 7 → 32 → 32 → 1 ReLU form:

* inputs ``S_xx S_yy S_zz S_xy S_xz S_yz |S|`` of the resolved strain rate, normalised by the
  training set's mean / std (``set_feature_stats``, model.py:66-80);
* inference runs **inside** the fused collision (``FusedSolver3D(closure=model)``): the strain
  comes from the non-equilibrium stress in registers and the two hidden layers run on the
  matrix cores (fp16 operands, fp32 accumulation), so the closure adds no HBM traffic;
* training runs on the GPU through the library's own forward / backward / Adam kernels
  (``tlbm_mlp_grad`` + ``tlbm_mlp_adam``), deterministic for a given seed;
* training data come from a DNS: block-averaged moments (``coarsen_moments``, the 3-D
  ``_coarsen_mean`` of ai/pipeline.py:93-107) → resolved strain and the sub-grid viscosity
  ``ν = max(0, −τ_ij S_ij / (2 S_kl S_kl))`` (``sgs_samples``).

Every op here calls ``libtlbm_hip.so``; torch only holds device memory, draws minibatch
permutations and computes validation statistics.
"""

from __future__ import annotations

import ctypes
import math
import time
from dataclasses import asdict, dataclass

import torch

from . import _native as nat

__all__ = [
    "EddyViscosityMLP3D",
    "TrainConfig3D",
    "train_eddy_viscosity_model_3d",
    "mlp_loss_and_grad",
    "coarsen_moments",
    "sgs_samples",
    "les_samples_from_populations",
    "les_tau",
    "run_dns_reference_3d",
    "les_initial_populations",
    "sizes",
]

N_IN = 7
N_HID = 32


def sizes() -> dict:
    """Parameter-vector and packed-blob sizes reported by the library."""
    v = [ctypes.c_int64() for _ in range(4)]
    nat.check(nat.lib().tlbm_mlp_sizes(*[ctypes.byref(x) for x in v]), "tlbm_mlp_sizes")
    return {"nparam": v[0].value, "ntotal": v[1].value, "packed_bytes": v[2].value, "batch_block": v[3].value}


def _linear_init(gen: torch.Generator, fan_out: int, fan_in: int) -> tuple[torch.Tensor, torch.Tensor]:
    """torch.nn.Linear's default init: U(−1/√fan_in, 1/√fan_in) for weight and bias."""
    bound = 1.0 / math.sqrt(fan_in)
    w = (torch.rand((fan_out, fan_in), generator=gen, dtype=torch.float64) * 2.0 - 1.0) * bound
    b = (torch.rand((fan_out,), generator=gen, dtype=torch.float64) * 2.0 - 1.0) * bound
    return w, b


class EddyViscosityMLP3D:
    """7 → 32 → 32 → 1 eddy-viscosity network held as one fp32 device vector.

    ``params`` layout: ``W1[32][7] b1[32] W2[32][32] b2[32] w3[32] b3 | mean[7] inv_std[7]``.
    ``model(x)`` returns ``(N, 1)`` ν_t ≥ 0 like the reference's ``EddyViscosityMLP.forward``.

    ``operands`` picks the matrix-core form the fused collision evaluates: ``"fp16"`` (fp16
    weights, inputs and layer-1 bias as two-term fp16 splits, fp16 activations) or ``"bf16"``
    (bf16 weights, inputs, bias and layer-1 activations as two-term bf16 splits); both
    accumulate in fp32.  Training and ``forward`` are fp32 either way.
    """

    in_features = N_IN
    hidden_features = N_HID

    def __init__(self, device: str | torch.device = "cuda:0", *, seed: int = 0,
                 params: torch.Tensor | None = None, operands: str = "fp16") -> None:
        if operands not in ("fp16", "bf16"):
            raise ValueError(f"operands must be 'fp16' or 'bf16', got {operands!r}")
        self.operands = operands
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError(f"EddyViscosityMLP3D lives on a GPU device, got {self.device}")
        sz = sizes()
        self.nparam, self.ntotal = sz["nparam"], sz["ntotal"]
        if params is not None:
            if params.numel() != self.ntotal:
                raise ValueError(f"parameter vector needs {self.ntotal} floats, got {params.numel()}")
            vec = params.detach().to(torch.float32).reshape(-1).cpu()
        else:
            gen = torch.Generator().manual_seed(int(seed))
            w1, b1 = _linear_init(gen, N_HID, N_IN)
            w2, b2 = _linear_init(gen, N_HID, N_HID)
            w3, b3 = _linear_init(gen, 1, N_HID)
            vec = torch.cat([w1.reshape(-1), b1, w2.reshape(-1), b2, w3.reshape(-1), b3,
                             torch.zeros(N_IN, dtype=torch.float64), torch.ones(N_IN, dtype=torch.float64)])
            vec = vec.to(torch.float32)
        self.params = vec.to(self.device).contiguous()
        self._packed = torch.empty(sz["packed_bytes"], dtype=torch.uint8, device=self.device)
        self._packed_version = -1
        self._version = 0

    # -- parameters ---------------------------------------------------------
    def mark_updated(self) -> None:
        """Call after writing ``params`` in place (the packed blob is rebuilt lazily)."""
        self._version += 1

    def set_feature_stats(self, mean: torch.Tensor, std: torch.Tensor) -> None:
        """Normalisation statistics (model.py:66-80: std clamped at 1e-6)."""
        mean = torch.as_tensor(mean, dtype=torch.float32).reshape(-1)
        std = torch.as_tensor(std, dtype=torch.float32).reshape(-1)
        if mean.shape != (N_IN,) or std.shape != (N_IN,):
            raise ValueError(f"Normalization stats must have shape ({N_IN},), got {tuple(mean.shape)} "
                             f"and {tuple(std.shape)}")
        self.params[self.nparam:self.nparam + N_IN] = mean.to(self.device)
        self.params[self.nparam + N_IN:] = 1.0 / std.clamp_min(1e-6).to(self.device)
        self.mark_updated()

    @property
    def tau_mode(self) -> int:
        """The fused collision's τ mode for this network's operand form."""
        return nat.TAU_MLP_BF16 if self.operands == "bf16" else nat.TAU_MLP

    @property
    def packed(self) -> torch.Tensor:
        """Matrix-core operand blob (fp16 or bf16 weights) for the fused collision, rebuilt when stale."""
        if self._packed_version != self._version:
            lib = nat.lib()
            pack = lib.tlbm_mlp_pack_bf16 if self.operands == "bf16" else lib.tlbm_mlp_pack
            nat.check(pack(self.params.data_ptr(), self._packed.data_ptr(), nat.stream_of(self.params)),
                      "tlbm_mlp_pack")
            self._packed_version = self._version
        return self._packed

    def state_dict(self) -> dict:
        return {"params": self.params.detach().cpu().clone(), "arch": {"in_features": N_IN, "hidden_features": N_HID,
                                                                        "n_hidden_layers": 2, "activation": "relu"}}

    def load_state_dict(self, sd: dict) -> None:
        p = sd["params"]
        if p.numel() != self.ntotal:
            raise ValueError(f"state dict holds {p.numel()} parameters, expected {self.ntotal}")
        self.params.copy_(p.reshape(-1).to(self.device, torch.float32))
        self.mark_updated()

    # -- inference ------------------------------------------------------------
    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return self.forward(x)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """fp32 ν_t for features ``x`` of shape (N, 7) → (N, 1)."""
        nat.require_gpu(x, "features")
        if x.ndim != 2 or x.shape[-1] != N_IN:
            raise ValueError(f"Expected input with {N_IN} features, got tensor of shape {tuple(x.shape)}")
        x = x.to(device=self.device, dtype=torch.float32).contiguous()
        y = torch.empty((x.shape[0], 1), dtype=torch.float32, device=self.device)
        nat.check(nat.lib().tlbm_mlp_predict(self.params.data_ptr(), x.data_ptr(), int(x.shape[0]), y.data_ptr(),
                                             nat.stream_of(x)), "tlbm_mlp_predict")
        return y


def mlp_loss_and_grad(model: EddyViscosityMLP3D, x: torch.Tensor, t: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """(MSE loss, gradient w.r.t. the nparam trainable parameters) on the GPU."""
    x = x.to(device=model.device, dtype=torch.float32).contiguous()
    t = t.to(device=model.device, dtype=torch.float32).reshape(-1).contiguous()
    if x.ndim != 2 or x.shape[1] != N_IN or x.shape[0] != t.shape[0]:
        raise ValueError("x must be (N, 7) and t (N,) or (N, 1)")
    bb = sizes()["batch_block"]
    nblk = (x.shape[0] + bb - 1) // bb
    partials = torch.empty((nblk, model.nparam + 1), dtype=torch.float32, device=model.device)
    loss = torch.empty(1, dtype=torch.float32, device=model.device)
    grad = torch.empty(model.nparam, dtype=torch.float32, device=model.device)
    lib, s = nat.lib(), nat.stream_of(x)
    nat.check(lib.tlbm_mlp_grad(model.params.data_ptr(), x.data_ptr(), t.data_ptr(), int(x.shape[0]),
                                partials.data_ptr(), s), "tlbm_mlp_grad")
    nat.check(lib.tlbm_mlp_adam(model.params.data_ptr(), partials.data_ptr(), int(nblk), None, None, 1,
                                0.0, 0.9, 0.999, 1e-8, 0.0, loss.data_ptr(), grad.data_ptr(), s), "tlbm_mlp_adam")
    return loss[0], grad


class _Adam:
    """Adam state on the device; one ``step`` = gradient kernel + reduce/clip/update kernel."""

    def __init__(self, model: EddyViscosityMLP3D, lr: float, betas=(0.9, 0.999), eps: float = 1e-8,
                 clip_norm: float | None = 1.0, max_batch: int = 4096) -> None:
        self.model = model
        self.lr, self.b1, self.b2, self.eps = float(lr), float(betas[0]), float(betas[1]), float(eps)
        self.clip = float(clip_norm) if clip_norm is not None else 0.0
        self.m = torch.zeros(model.nparam, dtype=torch.float32, device=model.device)
        self.v = torch.zeros_like(self.m)
        self.t = 0
        self.bb = sizes()["batch_block"]
        self.partials = torch.empty(((max_batch + self.bb - 1) // self.bb, model.nparam + 1), dtype=torch.float32,
                                    device=model.device)
        self.loss = torch.zeros(1, dtype=torch.float32, device=model.device)

    def step(self, x: torch.Tensor, t: torch.Tensor) -> torch.Tensor:
        n = int(x.shape[0])
        nblk = (n + self.bb - 1) // self.bb
        if nblk > self.partials.shape[0]:
            raise ValueError(f"minibatch of {n} exceeds the optimiser's buffer")
        self.t += 1
        lib, s, p = nat.lib(), nat.stream_of(x), self.model.params.data_ptr()
        nat.check(lib.tlbm_mlp_grad(p, x.data_ptr(), t.data_ptr(), n, self.partials.data_ptr(), s), "tlbm_mlp_grad")
        nat.check(lib.tlbm_mlp_adam(p, self.partials.data_ptr(), nblk, self.m.data_ptr(), self.v.data_ptr(), self.t,
                                    self.lr, self.b1, self.b2, self.eps, self.clip, self.loss.data_ptr(), None, s),
                  "tlbm_mlp_adam")
        self.model.mark_updated()
        return self.loss


@dataclass
class TrainConfig3D:
    """Hyper-parameters of :func:`train_eddy_viscosity_model_3d` (defaults of ai/train.py:31-46)."""

    epochs: int = 20
    batch_size: int = 4096
    learning_rate: float = 1e-3
    val_fraction: float = 0.1
    seed: int = 0
    patience: int | None = None
    gradient_clip_norm: float | None = 1.0


def _r2(pred: torch.Tensor, y: torch.Tensor) -> float:
    ss_res = torch.sum((pred - y) ** 2)
    ss_tot = torch.sum((y - y.mean()) ** 2).clamp_min(1e-30)
    return float(1.0 - ss_res / ss_tot)


def train_eddy_viscosity_model_3d(features: torch.Tensor, targets: torch.Tensor, config: TrainConfig3D | None = None,
                                  *, device: str | torch.device = "cuda:0",
                                  model: EddyViscosityMLP3D | None = None, log=None) -> tuple[EddyViscosityMLP3D, dict]:
    """Fit an :class:`EddyViscosityMLP3D` to (features (N, 7), ν targets (N,)) on the GPU.

    Follows ``train_eddy_viscosity_model`` (ai/train.py:310-464): seeded train/validation split,
    feature mean / std from the training part, shuffled minibatches from a CPU generator, MSE,
    gradient-norm clipping, Adam, and the best-validation parameters kept.  ``log(msg)``, when
    given, receives one line per epoch.
    """
    cfg = config or TrainConfig3D()
    device = torch.device(device) if model is None else model.device
    x = features.to(device=device, dtype=torch.float32).reshape(-1, N_IN).contiguous()
    y = targets.to(device=device, dtype=torch.float32).reshape(-1).contiguous()
    n = int(x.shape[0])
    if n < 4:
        raise ValueError(f"Dataset is too small to train ({n} samples).")
    gen = torch.Generator().manual_seed(int(cfg.seed))
    perm = torch.randperm(n, generator=gen).to(device)
    n_val = max(1, int(round(n * float(cfg.val_fraction))))
    val_idx, tr_idx = perm[:n_val], perm[n_val:]
    x_tr, y_tr = x.index_select(0, tr_idx).contiguous(), y.index_select(0, tr_idx).contiguous()
    x_val, y_val = x.index_select(0, val_idx).contiguous(), y.index_select(0, val_idx).contiguous()
    if model is None:
        model = EddyViscosityMLP3D(device, seed=int(cfg.seed))
    model.set_feature_stats(x_tr.mean(dim=0).cpu(), x_tr.std(dim=0, unbiased=False).cpu())
    n_tr = int(x_tr.shape[0])
    bs = max(1, min(int(cfg.batch_size), n_tr))
    opt = _Adam(model, cfg.learning_rate, clip_norm=cfg.gradient_clip_norm, max_batch=bs)
    history: list[dict] = []
    best = (float("inf"), model.params.clone(), 0)
    stale = 0
    t0 = time.perf_counter()
    for epoch in range(int(cfg.epochs)):
        order = torch.randperm(n_tr, generator=gen).to(device)
        ep_loss = torch.zeros((), dtype=torch.float64, device=device)
        for s in range(0, n_tr, bs):
            idx = order[s:s + bs]
            xb = x_tr.index_select(0, idx).contiguous()
            yb = y_tr.index_select(0, idx).contiguous()
            ep_loss += opt.step(xb, yb)[0].double() * xb.shape[0]
        pred = model(x_val).reshape(-1)
        val_mse = float(torch.mean((pred - y_val) ** 2))
        met = {"epoch": epoch, "train_mse": float(ep_loss) / n_tr, "val_mse": val_mse,
               "val_mae": float(torch.mean(torch.abs(pred - y_val))), "val_r2": _r2(pred, y_val)}
        history.append(met)
        if log is not None:
            log(f"epoch {epoch}: train mse {met['train_mse']:.3e}, val mse {val_mse:.3e}, val R2 {met['val_r2']:.3f}")
        if val_mse < best[0] - 1e-30:
            best = (val_mse, model.params.clone(), epoch)
            stale = 0
        else:
            stale += 1
        if cfg.patience is not None and stale > int(cfg.patience):
            break
    model.params.copy_(best[1])
    model.mark_updated()
    final = history[best[2]]
    meta = {"arch": {"in_features": N_IN, "hidden_features": N_HID, "n_hidden_layers": 2, "activation": "relu"},
            "config": asdict(cfg), "backend": "hip", "n_samples_train": n_tr, "n_samples_val": int(x_val.shape[0]),
            "history": history, "best_epoch": best[2], "final_train_mse": final["train_mse"],
            "final_val_mse": final["val_mse"], "final_val_mae": final["val_mae"], "final_val_r2": final["val_r2"],
            "stopped_early": len(history) < int(cfg.epochs), "training_time_s": time.perf_counter() - t0}
    return model, meta


# ---------------------------------------------------------------------------
# DNS -> LES training data
# ---------------------------------------------------------------------------


def coarsen_moments(f: torch.Tensor, factor: int, *, deviation: bool = False) -> torch.Tensor:
    """Block averages over ``factor``³ cells of ρ, u and u_i u_j from populations (Q, nz, ny, nx).

    Returns (10, nz/r, ny/r, nx/r) float32: ρ, u_x, u_y, u_z, <u_x u_x>, <u_y u_y>, <u_z u_z>,
    <u_x u_y>, <u_x u_z>, <u_y u_z> (velocities are per fine cell, averaged in FP64).
    """
    nat.require_gpu(f, "f")
    q, nz, ny, nx = nat.grid_of(f)
    if q == 9:
        raise NotImplementedError("coarse-graining is built for the 3-D lattices")
    r = int(factor)
    if r < 1 or nz % r or ny % r or nx % r:
        raise ValueError(f"shapes must match: grid {(nz, ny, nx)} is not a multiple of factor {r}")
    out = torch.empty((10, nz // r, ny // r, nx // r), dtype=torch.float32, device=f.device)
    st = nat.strides_of(f)
    nat.check(nat.lib().tlbm_coarsen_moments(q, nat.store_code(f.dtype, deviation), f.data_ptr(), st, nz, ny, nx, r,
                                             out.data_ptr(), nat.stream_of(f)), "coarsen_moments")
    return out


def sgs_samples(fields: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """(features (N, 7), ν labels (N,)) from coarse moments (10, nz, ny, nx) (periodic differences)."""
    nat.require_gpu(fields, "fields")
    if fields.ndim != 4 or fields.shape[0] != 10 or fields.dtype != torch.float32:
        raise ValueError("fields must be the (10, nz, ny, nx) float32 output of coarsen_moments")
    fields = fields.contiguous()
    _, nz, ny, nx = (int(s) for s in fields.shape)
    n = nz * ny * nx
    feats = torch.empty((n, N_IN), dtype=torch.float32, device=fields.device)
    labels = torch.empty(n, dtype=torch.float32, device=fields.device)
    nat.check(nat.lib().tlbm_sgs_samples(fields.data_ptr(), nz, ny, nx, feats.data_ptr(), labels.data_ptr(),
                                         nat.stream_of(fields)), "sgs_samples")
    return feats, labels


def les_samples_from_populations(f: torch.Tensor, factor: int, *, deviation: bool = False):
    """``sgs_samples(coarsen_moments(f, factor))``."""
    return sgs_samples(coarsen_moments(f, factor, deviation=deviation))


def les_tau(tau_dns: float, dns_scale: int) -> float:
    """Relaxation time of the coarse grid under acoustic scaling (Δx, Δt both × r): ν_les = ν_dns / r."""
    return 3.0 * ((float(tau_dns) - 0.5) / 3.0) / int(dns_scale) + 0.5


def run_dns_reference_3d(n_les: int, tau_dns: float, n_steps_les: int, sample_every: int, seed: int,
                         device: str | torch.device = "cuda:0", *, dns_scale: int = 3, k0: float = 4.0,
                         u_rms: float = 0.03) -> list[torch.Tensor]:
    """Coarse moment snapshots of a fine BGK run (the 3-D form of ``_run_dns_reference``,
    ai/pipeline.py:110-149).

    The fine grid is ``n_les · dns_scale`` per edge, started at rest density from the
    random solenoidal field (benchmarks/pending/dit_turbulence/run.py:95-125, spectrum peak ``k0``,
    ``u_rms``), advanced ``dns_scale`` fine steps per coarse step and block-averaged every
    ``sample_every`` coarse steps.  Returns ``coarsen_moments`` fields, the first at t = 0.
    """
    from . import fields, ops
    from .fused import FusedSolver3D

    r = int(dns_scale)
    n = int(n_les) * r
    u = fields.random_solenoidal_field(n, k0, u_rms, seed=seed, device=device, dtype=torch.float32)
    f = ops.equilibrium3d(torch.ones((n, n, n), device=device), u[0].contiguous(), u[1].contiguous(),
                          u[2].contiguous())
    del u
    solver = FusedSolver3D(n, n, n, tau_dns, device, torch.float32, layout="contiguous")
    snaps = [coarsen_moments(f, r)]
    for k in range(1, int(n_steps_les) + 1):
        f, _ = solver.run(f, r)  # two fine steps per launch where the solver builds that case
        if sample_every > 0 and k % sample_every == 0:
            snaps.append(coarsen_moments(f, r))
    return snaps


def les_initial_populations(fields: torch.Tensor) -> torch.Tensor:
    """Equilibrium populations of a coarse snapshot's ρ̄ and ū (for starting an LES from filtered DNS)."""
    from . import ops

    return ops.equilibrium3d(fields[0].contiguous(), fields[1].contiguous(), fields[2].contiguous(),
                             fields[3].contiguous())


def clone_model(model: EddyViscosityMLP3D) -> EddyViscosityMLP3D:
    return EddyViscosityMLP3D(model.device, params=model.params, operands=model.operands)
