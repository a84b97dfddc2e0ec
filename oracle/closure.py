"""Learned eddy-viscosity closure: CPU restatement (test infrastructure only).

* The network mirrors the reference's ``EddyViscosityMLP``
  (src/tensorlbm/ai/model.py:40-92): feature normalisation, hidden Linear
  layers, a scalar Linear output and Softplus(β=10) so ν_t ≥ 0.  The 3-D form
  used here is 7 → 32 → 32 → 1 with ReLU hidden activations.  It has no
  reference counterpart (the reference's network is 2-D, 3 features).
* Training is MSE with Adam (ai/train.py:382-435 uses the same loss/optimiser).
* Data: block-averaged DNS fields (ai/pipeline.py:93-107 ``_coarsen_mean``,
  3-D by factor r), strain from periodic central differences
  (ai/dataset.py:41-72 in 3-D), labels from the sub-grid stress
  τ_ij = <u_i u_j> − <u_i><u_j> projected on the resolved strain,
  ν_t = max(0, −τ_ij S_ij / (2 S_kl S_kl)).
* Initial field: random-phase solenoidal velocity with E(k) ∝ k⁴ exp(−2k²/k0²)
  (benchmarks/pending/dit_turbulence/run.py:95-125).

Pinned in ``tests/test_closure_oracle.py``: gradients by torch autograd and
finite differences, Adam + clipping by ``torch.optim.Adam`` and
``clip_grad_norm_``, ``fp16`` by torch's half cast, ``coarsen`` / ``strain``
/ ``sgs_samples`` by constructed fields with closed-form answers, and the
initial field by divergence / spectrum-peak / u_rms properties.
"""

from __future__ import annotations

import numpy as np

N_IN, N_HID = 7, 32
BETA = 10.0
PARAM_SHAPES = (("w1", (N_HID, N_IN)), ("b1", (N_HID,)), ("w2", (N_HID, N_HID)), ("b2", (N_HID,)),
                ("w3", (N_HID,)), ("b3", (1,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in PARAM_SHAPES)  # 1345


def unpack(vec: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in PARAM_SHAPES:
        n = int(np.prod(shape))
        out[name] = vec[o : o + n].reshape(shape)
        o += n
    return out


def pack(p: dict) -> np.ndarray:
    return np.concatenate([np.asarray(p[name], dtype=np.float64).reshape(-1) for name, _ in PARAM_SHAPES])


def softplus(z, beta=BETA):
    bz = beta * z
    return np.where(bz > 20.0, z, np.log1p(np.exp(np.minimum(bz, 20.0))) / beta)


def forward(vec, x, mean, inv_std):
    """ν_t for features x (N, 7); returns (y, cache)."""
    p = unpack(np.asarray(vec, dtype=np.float64))
    xh = (np.asarray(x, np.float64) - mean) * inv_std
    z1 = xh @ p["w1"].T + p["b1"]
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ p["w2"].T + p["b2"]
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ p["w3"] + p["b3"][0]
    return softplus(z3), (xh, z1, h1, z2, h2, z3, p)


def loss_and_grad(vec, x, t, mean, inv_std):
    """MSE loss mean((y − t)²) and its gradient with respect to the packed parameters."""
    y, (xh, z1, h1, z2, h2, z3, p) = forward(vec, x, mean, inv_std)
    n = y.shape[0]
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    bz = BETA * z3
    # softplus' = sigmoid(β z), and 1 where the forward pass returns z itself (β z > 20)
    dz3 = dy * np.where(bz > 20.0, 1.0, 1.0 / (1.0 + np.exp(-np.clip(bz, -60.0, 60.0))))
    g = {"w3": h2.T @ dz3, "b3": np.array([dz3.sum()])}
    dh2 = np.outer(dz3, p["w3"])
    dz2 = dh2 * (z2 > 0)
    g["w2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dh1 = dz2 @ p["w2"]
    dz1 = dh1 * (z1 > 0)
    g["w1"] = dz1.T @ xh
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def clip_grad_norm(grad, max_norm):
    """torch.nn.utils.clip_grad_norm_ rule (ai/train.py:401-405): scale by min(1, max/(|g| + 1e-6))."""
    if max_norm is None or max_norm <= 0:
        return grad
    norm = float(np.sqrt(np.sum(np.asarray(grad, np.float64) ** 2)))
    return grad * min(1.0, max_norm / (norm + 1e-6))


def fp16(a):
    """Round fp32 values to binary16 (nearest, ties to even; overflow to ±inf) and return as float64."""
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float64)


def split_fp16(a):
    """Two-term fp16 split a ≈ hi + lo, hi = fp16(a), lo = fp16(a − hi), of fp32 values."""
    a32 = np.asarray(a, np.float32)
    hi = fp16(a32)
    lo = fp16((a32.astype(np.float64) - hi).astype(np.float32))
    return hi, lo


def bf16(a):
    """Round fp32 values to bfloat16 (nearest, ties to even; finite inputs) and return as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def split_bf16(a):
    """Two-term bf16 split a ≈ hi + lo, hi = bf16(a), lo = bf16(a − hi), of fp32 values."""
    a32 = np.asarray(a, np.float32)
    hi = bf16(a32)
    lo = bf16((a32.astype(np.float64) - hi).astype(np.float32))
    return hi, lo


def forward_mfma_bf16(vec, x, mean, inv_std):
    """ν_t of the bf16-operand form: bf16 W1 and W2; normalised inputs, the layer-1 bias and the
    layer-1 activations relu(z1) as two-term bf16 splits; wide accumulation; fp32 output layer."""
    p = unpack(np.asarray(vec, dtype=np.float64))
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731
    xh = f32((f32(x) - f32(mean)) * f32(inv_std))
    xhi, xlo = split_bf16(xh)
    w1 = bf16(p["w1"])
    b1h, b1l = split_bf16(p["b1"])
    z1 = f32(xhi @ w1.T + xlo @ w1.T + (b1h + b1l))
    hh, hl = split_bf16(np.maximum(z1, 0.0))
    w2 = bf16(p["w2"])
    z2 = f32(hh @ w2.T + hl @ w2.T + f32(p["b2"]))
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ f32(p["w3"]) + f32(p["b3"])[0]
    return softplus(z3)


def forward_mfma(vec, x, mean, inv_std):
    """ν_t as the fused collision evaluates it on the matrix cores: fp16 W1 and W2, normalised
    inputs and the layer-1 bias as two-term fp16 splits, layer-1 activations relu(z1) rounded
    once to fp16, wide accumulation, fp32 output layer."""
    p = unpack(np.asarray(vec, dtype=np.float64))
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731
    xh = f32((f32(x) - f32(mean)) * f32(inv_std))
    xhi, xlo = split_fp16(xh)
    w1 = fp16(p["w1"])
    b1h, b1l = split_fp16(p["b1"])  # the bias enters the tile as hi + lo against a constant 1
    z1 = f32(xhi @ w1.T + xlo @ w1.T + (b1h + b1l))
    h1 = fp16(np.maximum(z1, 0.0))
    w2 = fp16(p["w2"])
    z2 = f32(h1 @ w2.T + f32(p["b2"]))
    h2 = np.maximum(z2, 0.0)
    z3 = h2 @ f32(p["w3"]) + f32(p["b3"])[0]
    return softplus(z3)


def adam(vec, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (torch.optim.Adam defaults); returns (vec, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1**step)
    vh = v / (1 - b2**step)
    return vec - lr * mh / (np.sqrt(vh) + eps), m, v


# ---------------------------------------------------------------------------
# features from the non-equilibrium stress (the fused kernel's inputs)
# ---------------------------------------------------------------------------


def features_from_pi(pi, rho, tau_ref):
    """S = −Π^neq / (2 ρ c_s² τ_ref) (6 components xx yy zz xy xz yz) and |S| = √(2 S:S)."""
    s = [-1.5 * p / (rho * tau_ref) for p in pi]
    mag = np.sqrt(2.0 * (s[0] ** 2 + s[1] ** 2 + s[2] ** 2 + 2.0 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2)))
    return np.stack(s + [mag], axis=-1)


# ---------------------------------------------------------------------------
# DNS -> LES data
# ---------------------------------------------------------------------------


def coarsen(field, r):
    """Block average of a 3-D field by integer factor r (ai/pipeline.py:93-107 in 3-D)."""
    nz, ny, nx = field.shape
    return field.reshape(nz // r, r, ny // r, r, nx // r, r).mean(axis=(1, 3, 5))


def strain(u):
    """Periodic central-difference strain of a (3, nz, ny, nx) velocity (ai/dataset.py:41-72 in 3-D)."""
    g = np.empty((3, 3) + u.shape[1:])
    axes = {0: 2, 1: 1, 2: 0}  # x is the last array axis
    for i in range(3):
        for j in range(3):
            ax = axes[j]
            g[i, j] = 0.5 * (np.roll(u[i], -1, axis=ax) - np.roll(u[i], 1, axis=ax))
    s = 0.5 * (g + g.transpose(1, 0, 2, 3, 4))
    return s


def sgs_samples(rho, u, r):
    """(features (N,7), labels (N,)) from a fine velocity field u (3, n, n, n) filtered by factor r."""
    ub = np.stack([coarsen(u[i], r) for i in range(3)])
    uu = {}
    for i in range(3):
        for j in range(i, 3):
            uu[(i, j)] = coarsen(u[i] * u[j], r) - ub[i] * ub[j]
    s = strain(ub)
    comps = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    ss = sum((1.0 if i == j else 2.0) * s[i, j] ** 2 for i, j in comps)
    ts = sum((1.0 if i == j else 2.0) * uu[(i, j)] * s[i, j] for i, j in comps)
    nu = np.maximum(0.0, -ts / (2.0 * np.maximum(ss, 1e-30)))
    mag = np.sqrt(2.0 * ss)
    feats = np.stack([s[i, j] for i, j in comps] + [mag], axis=-1).reshape(-1, 7)
    del rho
    return feats, nu.reshape(-1)


def random_solenoidal_field(n, k0, u_rms, seed):
    """Divergence-free random-phase field with E(k) ∝ k⁴ exp(−2k²/k0²), exact u_rms (run.py:95-125)."""
    rng = np.random.default_rng(seed)
    kx = np.fft.rfftfreq(n) * n
    ky = np.fft.fftfreq(n) * n
    kz = np.fft.fftfreq(n) * n
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    K2 = KX * KX + KY * KY + KZ * KZ
    K = np.sqrt(K2)
    spec = K**4 * np.exp(-2.0 * (K / k0) ** 2)
    amp = np.sqrt(np.where(K > 0, spec / (4.0 * np.pi * np.where(K2 > 0, K2, 1.0)), 0.0))
    uh = np.stack([(rng.standard_normal(K.shape) + 1j * rng.standard_normal(K.shape)) * amp for _ in range(3)])
    kdotu = KX * uh[0] + KY * uh[1] + KZ * uh[2]
    fac = np.where(K2 > 0, kdotu / np.where(K2 > 0, K2, 1.0), 0.0)
    for i, kc in enumerate((KX, KY, KZ)):
        uh[i] -= kc * fac
    uh[:, K == 0] = 0.0
    u = np.fft.irfftn(uh, s=(n, n, n), axes=(1, 2, 3))
    return u * (u_rms / np.sqrt(np.mean(u**2)))  # the reference then casts to float32


def spectral_divergence(u):
    """max |k·û| / max |k||û| of a periodic field over the non-Nyquist modes (0 if solenoidal).

    Nyquist planes are left out: their sign of k is ambiguous on an even grid, and
    the real transform folds the two signs together."""
    n = u.shape[1]
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    kx = np.fft.rfftfreq(n) * n
    ky = np.fft.fftfreq(n) * n
    kz = np.fft.fftfreq(n) * n
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    keep = (np.abs(KX) < n / 2) & (np.abs(KY) < n / 2) & (np.abs(KZ) < n / 2)
    div = np.abs(np.where(keep, KX * uh[0] + KY * uh[1] + KZ * uh[2], 0.0)).max()
    scale = (np.sqrt(KX**2 + KY**2 + KZ**2) * np.sqrt((np.abs(uh) ** 2).sum(0))).max()
    return float(div / scale)


def shell_spectrum(u):
    """Shell-summed energy spectrum E(k), k = 0..n/2."""
    n = u.shape[1]
    uh = np.fft.fftn(u, axes=(1, 2, 3)) / n**3
    e = 0.5 * (np.abs(uh) ** 2).sum(0)
    k = np.fft.fftfreq(n) * n
    KZ, KY, KX = np.meshgrid(k, k, k, indexing="ij")
    kk = np.rint(np.sqrt(KX**2 + KY**2 + KZ**2)).astype(int)
    return np.bincount(kk.ravel(), weights=e.ravel())[: n // 2 + 1]


# ---------------------------------------------------------------------------
# the fused step with the learned closure
# ---------------------------------------------------------------------------


def fused_mlp_step(f, tau, params, cs_delta_sq=0.0, emulate_mfma=True, operands="fp16"):
    """Periodic pull-stream then BGK with τ_eff = max(τ + 3 ν_t, 0.501) (ai/inference.py:56-92).

    ν_t = MLP(S, |S|) with S = −Π^neq / (2 ρ c_s² τ_ref) (turbulence.py:1449-1480), τ_ref the
    clamped Smagorinsky τ of ``cs_delta_sq`` (τ itself when it is 0).  ``params`` is the full
    vector (trainable + mean + inv_std).  ``operands`` ("fp16" or "bf16") picks the matrix-core
    form that ``emulate_mfma`` reproduces.  Returns (f_out, ν_t).
    """
    from . import ref

    return mlp_collide(ref.stream(f), tau, params, cs_delta_sq, emulate_mfma, operands)


def mlp_collide(fin, tau, params, cs_delta_sq=0.0, emulate_mfma=True, operands="fp16"):
    """The collision half of :func:`fused_mlp_step`, on post-stream populations ``fin`` (after any
    boundary overrides and bounce-back, as the fused step orders them).  Returns (f_out, ν_t)."""
    from . import ref

    params = np.asarray(params, np.float64)
    vec, mean, istd = params[:N_PARAMS], params[N_PARAMS:N_PARAMS + N_IN], params[N_PARAMS + N_IN:]
    feq, rho_safe, _ = ref._feq_safe(fin)
    fneq = fin - feq
    pi = ref.neq_stress(fneq)
    te_ref = ref.smagorinsky_tau(tau, ref.pi_norm(fneq), rho_safe, cs_delta_sq)
    x = features_from_pi(pi, rho_safe, te_ref).reshape(-1, N_IN)
    mfma = forward_mfma_bf16 if operands == "bf16" else forward_mfma
    nu = mfma(vec, x, mean, istd) if emulate_mfma else forward(vec, x, mean, istd)[0]
    nu = nu.reshape(rho_safe.shape)
    te = np.maximum(tau + 3.0 * nu, 0.501)
    return fin - fneq / te[None], nu
