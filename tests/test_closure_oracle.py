"""Pins of the closure oracle (CPU): each check is independent of the formula it tests."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import closure as oc


def _params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, scale, oc.N_PARAMS)


def _data(n=50, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.normal(0.0, 1.0, (n, 7))
    t = np.abs(rng.normal(0.05, 0.03, n))
    mean = rng.normal(0.0, 0.1, 7)
    istd = 1.0 / rng.uniform(0.5, 2.0, 7)
    return x, t, mean, istd


def _torch_net(vec):
    p = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in oc.unpack(vec).items()}
    return p


def _torch_forward(p, x, mean, istd):
    xh = (torch.tensor(x) - torch.tensor(mean)) * torch.tensor(istd)
    h1 = torch.relu(xh @ p["w1"].T + p["b1"])
    h2 = torch.relu(h1 @ p["w2"].T + p["b2"])
    z3 = h2 @ p["w3"] + p["b3"][0]
    return torch.nn.functional.softplus(z3, beta=10.0, threshold=20.0)


def test_forward_matches_torch_modules():
    vec = _params()
    x, _, mean, istd = _data()
    y, _ = oc.forward(vec, x, mean, istd)
    ref = _torch_forward(_torch_net(vec), x, mean, istd).detach().numpy()
    np.testing.assert_allclose(y, ref, rtol=1e-13, atol=1e-15)
    assert (y >= 0).all()


def test_gradient_matches_torch_autograd():
    vec = _params(3)
    x, t, mean, istd = _data(64, 4)
    loss, g = oc.loss_and_grad(vec, x, t, mean, istd)
    p = _torch_net(vec)
    y = _torch_forward(p, x, mean, istd)
    tl = torch.mean((y - torch.tensor(t)) ** 2)
    tl.backward()
    ref = oc.pack({k: v.grad.numpy() for k, v in p.items()})
    assert abs(loss - float(tl.detach())) < 1e-14
    np.testing.assert_allclose(g, ref, rtol=1e-10, atol=1e-14)


def test_gradient_matches_finite_differences():
    vec = _params(5)
    x, t, mean, istd = _data(40, 6)
    _, g = oc.loss_and_grad(vec, x, t, mean, istd)
    rng = np.random.default_rng(7)
    for k in rng.choice(oc.N_PARAMS, 25, replace=False):
        h = 1e-6
        vp, vm = vec.copy(), vec.copy()
        vp[k] += h
        vm[k] -= h
        fd = (oc.loss_and_grad(vp, x, t, mean, istd)[0] - oc.loss_and_grad(vm, x, t, mean, istd)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_adam_and_clipping_match_torch_optim():
    vec = _params(8)
    x, t, mean, istd = _data(32, 9)
    p = torch.tensor(vec, requires_grad=True)
    opt = torch.optim.Adam([p], lr=1e-2)
    m = np.zeros_like(vec)
    v = np.zeros_like(vec)
    cur = vec.copy()
    for step in range(1, 6):
        _, g = oc.loss_and_grad(cur, x, t, mean, istd)
        gc = oc.clip_grad_norm(g, 0.05)
        cur, m, v = oc.adam(cur, gc, m, v, step, lr=1e-2)
        opt.zero_grad()
        p.grad = torch.tensor(g.copy())
        torch.nn.utils.clip_grad_norm_([p], max_norm=0.05)
        np.testing.assert_allclose(p.grad.numpy(), gc, rtol=1e-12)
        opt.step()
        with torch.no_grad():
            np.testing.assert_allclose(cur, p.numpy(), rtol=1e-10, atol=1e-12)
        p.data = torch.tensor(cur)  # same parameters for the next gradient


def test_fp16_rounding_matches_torch():
    rng = np.random.default_rng(11)
    a = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 5000), rng.normal(0, 3e4, 500),
                        [0.0, -0.0, 1.0, 1.0 + 2.0**-11, 1.0 + 3 * 2.0**-11, 7e4, 2.0**-24, 2.0**-26]])
    ref = torch.tensor(a, dtype=torch.float32).to(torch.float16).double().numpy()
    np.testing.assert_array_equal(oc.fp16(a), ref)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(12)
    a = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-3, 5000), rng.normal(0, 3e30, 500),
                        [0.0, -0.0, 1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8, 1.0 + 2.0**-9, 2.0**-130, -7.5e-5]])
    ref = torch.tensor(a, dtype=torch.float32).to(torch.bfloat16).double().numpy()
    np.testing.assert_array_equal(oc.bf16(a), ref)


def test_bf16_split_carries_sixteen_bits():
    rng = np.random.default_rng(13)
    a = rng.normal(0, 1, 20000).astype(np.float32)
    hi, lo = oc.split_bf16(a)
    err = np.abs(hi + lo - a.astype(np.float64))
    assert np.all(err <= 2.0**-16 * np.abs(a) + 1e-38)
    assert np.abs(hi - a).max() > 2.0**-10  # hi alone is bf16


def test_forward_mfma_bf16_is_close_to_fp64_forward():
    """bf16 weights move ν_t by a fixed perturbation of the network (~2^-9 per weight)."""
    vec = _params(12, 0.2)
    x, _, mean, istd = _data(200, 13)
    y, _ = oc.forward(vec, x, mean, istd)
    yb = oc.forward_mfma_bf16(vec, x, mean, istd)
    assert np.max(np.abs(yb - y)) < 0.03 * np.max(np.abs(y)) + 1e-6
    assert np.max(np.abs(yb - y)) > np.max(np.abs(oc.forward_mfma(vec, x, mean, istd) - y))


def test_forward_mfma_bf16_is_smooth_in_its_inputs():
    """The two-term splits keep ~16 bits of the inputs and of the layer-1 activations, so a 1e-6
    relative input nudge moves ν_t by accumulation error in almost every sample."""
    vec = _params(15, 0.3)
    x, _, mean, istd = _data(2000, 16)
    y0 = oc.forward_mfma_bf16(vec, x, mean, istd)
    y1 = oc.forward_mfma_bf16(vec, x * (1 + 1e-6), mean, istd)
    r = np.abs(y1 - y0) / np.maximum(np.abs(y0), 1e-3)
    print(f"bf16 form: 99th percentile {np.quantile(r, 0.99):.2e}, max {r.max():.2e}")
    # measured 5.7e-5 and 9.8e-5: smoother than the fp16 form, whose activations round once to 11 bits
    assert np.quantile(r, 0.99) < 1.5e-4
    assert r.max() < 5e-4


def test_forward_mfma_is_close_to_fp64_forward():
    vec = _params(12, 0.2)
    x, _, mean, istd = _data(200, 13)
    y, _ = oc.forward(vec, x, mean, istd)
    yb = oc.forward_mfma(vec, x, mean, istd)
    assert np.max(np.abs(yb - y)) < 0.01 * np.max(np.abs(y)) + 1e-6
    assert np.max(np.abs(yb - y)) > 0.0  # the emulation does round


# ---------------------------------------------------------------------------
# data pipeline
# ---------------------------------------------------------------------------


def test_coarsen_block_average():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(6, 9, 12))
    c = oc.coarsen(a, 3)
    assert c.shape == (2, 3, 4)
    assert abs(c[1, 2, 3] - a[3:6, 6:9, 9:12].mean()) < 1e-14
    assert abs(c.mean() - a.mean()) < 1e-14


def test_strain_of_a_shear_wave_is_the_central_difference_closed_form():
    n = 16
    k = 2 * np.pi / n
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u = np.zeros((3, n, n, n))
    u[0] = np.sin(k * y)  # u_x(y): S_xy = S_yx = ½ du_x/dy
    u[2] = np.cos(k * x)  # u_z(x): S_xz = ½ du_z/dx
    s = oc.strain(u)
    # periodic central difference of sin(k y) is cos(k y) sin(k)
    np.testing.assert_allclose(s[0, 1], 0.5 * np.cos(k * y) * np.sin(k), atol=1e-14)
    np.testing.assert_allclose(s[0, 2], -0.5 * np.sin(k * x) * np.sin(k), atol=1e-14)
    for i, j in ((0, 0), (1, 1), (2, 2), (1, 2)):
        assert np.abs(s[i, j]).max() < 1e-14
    np.testing.assert_allclose(s, s.transpose(1, 0, 2, 3, 4))


def test_sgs_samples_on_a_constructed_field():
    """Fine field = resolved shear + a sub-filter ±a wiggle in u_x along x: τ_xx = a², S = shear only."""
    r, nc = 2, 8
    n = r * nc
    k = 2 * np.pi / nc
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    yc = y // r
    u = np.zeros((3, n, n, n))
    amp, a = 0.01, 1e-3
    u[0] = amp * np.sin(k * yc) + a * np.where(x % 2 == 0, 1.0, -1.0)
    feats, nu = oc.sgs_samples(np.ones((n, n, n)), u, r)
    # resolved strain: only S_xy = ½ amp cos(k Y) sin(k); τ_xx = a² is not projected on S_xy
    sxy = 0.5 * amp * np.cos(k * np.arange(nc)) * np.sin(k)
    f = feats.reshape(nc, nc, nc, 7)
    np.testing.assert_allclose(f[:, :, :, 3], np.broadcast_to(sxy[None, :, None], (nc, nc, nc)), atol=1e-15)
    np.testing.assert_allclose(f[..., 6], np.sqrt(4.0 * f[..., 3] ** 2), atol=1e-15)
    assert np.abs(nu).max() < 1e-20
    # now make the wiggle correlate with the shear: u_y wiggle in phase with u_x's gives τ_xy = a²
    u[1] = a * np.where(x % 2 == 0, 1.0, -1.0)
    _, nu2 = oc.sgs_samples(np.ones((n, n, n)), u, r)
    ss = 2.0 * f[..., 3] ** 2
    expect = np.maximum(0.0, -(2.0 * a * a * f[..., 3]) / (2.0 * np.maximum(ss, 1e-30)))
    np.testing.assert_allclose(nu2.reshape(nc, nc, nc), expect, rtol=1e-9, atol=1e-18)


@pytest.mark.parametrize("n,k0", [(32, 4.0), (24, 3.0)])
def test_random_solenoidal_field_properties(n, k0):
    u = oc.random_solenoidal_field(n, k0, 0.03, seed=42)
    assert u.shape == (3, n, n, n)
    assert abs(np.sqrt(np.mean(u**2)) - 0.03) < 1e-12
    assert np.abs(u.mean(axis=(1, 2, 3))).max() < 1e-15
    assert oc.spectral_divergence(u) < 1e-12
    e = oc.shell_spectrum(u)
    assert abs(int(np.argmax(e)) - k0) <= 1  # E(k) ∝ k⁴ exp(−2k²/k0²) peaks at k0
    v = oc.random_solenoidal_field(n, k0, 0.03, seed=42)
    np.testing.assert_array_equal(u, v)


def test_features_from_pi_use_the_reference_denominator():
    """S = −Π/(2 ρ τ c_s²) as in _strain_from_fneq_3d (turbulence.py:1449-1480), |S| = √(2 S:S)."""
    pi = [np.array([0.3]), np.array([-0.1]), np.array([0.2]), np.array([0.05]), np.array([-0.02]), np.array([0.01])]
    f = oc.features_from_pi(pi, np.array([1.1]), 0.8)
    denom = 2.0 * 1.1 * 0.8 / 3.0
    np.testing.assert_allclose(f[0, :6], [-p[0] / denom for p in pi], rtol=1e-14)
    s = f[0, :6]
    np.testing.assert_allclose(f[0, 6], np.sqrt(2 * (s[0] ** 2 + s[1] ** 2 + s[2] ** 2 + 2 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2))))


def test_two_term_fp16_split_carries_22_bits():
    rng = np.random.default_rng(14)
    a = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 30, 4000), rng.normal(0, 1e-4, 4000)]).astype(np.float32)
    hi, lo = oc.split_fp16(a)
    err = np.abs(hi + lo - a.astype(np.float64))
    # 11 + 11 bits in the normal range; below it lo is an fp16 subnormal (step 2^-24)
    assert np.all(err <= np.maximum(2.0**-21 * np.abs(a), 2.0**-25))
    assert np.abs(hi - a).max() > 0  # hi alone is fp16


def test_forward_mfma_is_smooth_in_its_inputs():
    """Inputs carry ~22 bits and activations 11: a relative input nudge of 1e-6 leaves ν unchanged
    or moves it by accumulation error in almost every sample; where an activation's fp16 rounding
    flips, ν moves by less than one bf16 step of the output (2^-8 ≈ 4e-3 relative)."""
    vec = _params(15, 0.3)
    x, _, mean, istd = _data(2000, 16)
    y0 = oc.forward_mfma(vec, x, mean, istd)
    y1 = oc.forward_mfma(vec, x * (1 + 1e-6), mean, istd)
    r = np.abs(y1 - y0) / np.maximum(np.abs(y0), 1e-3)
    assert np.quantile(r, 0.99) < 2e-4
    assert r.max() < 4e-3