"""The LBMStepExecutor seam (lbm_step.py:70-686) on the HIP kernels."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref
from tensorlbm_amd.executor import LBMStepExecutor
from tensorlbm_amd.reporters import CallbackReporter
from tests import _fields as F

pytestmark = pytest.mark.gpu

SH = (6, 8, 10)


def _ex(dev, dtype=torch.float64, **kw):
    return LBMStepExecutor("D3Q19", device=dev, nx=SH[2], ny=SH[1], nz=SH[0], tau=0.7, dtype=dtype, **kw)


def _bgk_eager(f, tau):
    rho, ux, uy, uz = ref.macroscopic(f)
    return f - (f - ref.equilibrium(ref.D3Q19, rho, ux, uy, uz)) / tau


def test_internal_bgk_steps_match_oracle(dev):
    f0 = F.random_state(19, SH, seed=41, noise=1e-2)
    ex = _ex(dev, collide_fn="bgk")
    f, diag = ex.step(torch.tensor(f0, device=dev))
    want = ref.stream(_bgk_eager(f0, 0.7))
    np.testing.assert_allclose(f.cpu().numpy(), want, rtol=1e-12, atol=1e-14)
    assert diag == {}


def test_fused_run_equals_step_loop(dev):
    f0 = torch.tensor(F.random_state(19, SH, seed=42, noise=1e-2), device=dev)
    ex = _ex(dev, collide_fn="bgk")
    fast, diags = ex.run(f0, 7)
    assert len(diags) == 7
    slow = f0
    for _ in range(7):
        slow, _ = ex.step(slow)
    np.testing.assert_allclose(fast.cpu().numpy(), slow.cpu().numpy(), rtol=1e-13, atol=1e-15)
    want = F.random_state(19, SH, seed=42, noise=1e-2)
    for _ in range(7):
        want = ref.stream(_bgk_eager(want, 0.7))
    np.testing.assert_allclose(fast.cpu().numpy(), want, rtol=1e-12, atol=1e-14)


def test_callables_plug_in(dev):
    f0 = F.random_state(19, SH, seed=43, noise=1e-2)
    calls = []

    def boundary(f, scale=1.0):
        calls.append(scale)
        return f

    def force(f, rho, ux, uy, uz, k=0):
        return torch.stack([ux.sum(), uy.sum(), uz.sum()]) * k

    ex = _ex(dev, collide_fn=tl.collide_mrt3d, stream_fn=tl.stream3d, boundary_fn=boundary,
             boundary_kwargs={"scale": 2.0}, force_fn=force, force_kwargs={"k": 3})
    f, diag = ex.step(torch.tensor(f0, device=dev))
    want = ref.stream(ref.collide_mrt(f0, 0.7))
    np.testing.assert_allclose(f.cpu().numpy(), want, rtol=1e-12, atol=1e-14)
    assert calls == [2.0]
    _, ux, uy, uz = ref.macroscopic(want)
    np.testing.assert_allclose(diag["forces"].cpu().numpy(), 3 * np.array([ux.sum(), uy.sum(), uz.sum()]), rtol=1e-10)
    assert diag["max_speed"] > 0 and abs(diag["mean_rho"] - ref.macroscopic(want)[0].mean()) < 1e-12


def test_mass_correction_and_reporters(dev):
    f0 = torch.tensor(F.random_state(19, SH, seed=44, noise=1e-2), device=dev)
    target = float(f0.sum()) * 1.001
    seen = []
    rep = CallbackReporter(lambda ctx: seen.append((ctx.step, float(ctx.macroscopic()[0].sum()))), interval=2)
    ex = _ex(dev, collide_fn="bgk", target_mass=target)
    f, diags = ex.run(f0, 5, reporters=[rep])
    assert abs(float(f.sum()) - target) < 1e-10 * target
    assert [s for s, _ in seen] == [2, 4]
    assert all(abs(m - target) < 1e-9 * target for _, m in seen)


def test_unsupported_options_raise(dev):
    with pytest.raises(NotImplementedError):
        _ex(dev, collide_fn="bgk", wall_fn=True)
    with pytest.raises(ValueError):
        _ex(dev, collide_fn="trt")
    with pytest.raises(ValueError):
        LBMStepExecutor("D2Q9", collide_fn="bgk", device=dev, nx=4, ny=4, nz=1)
