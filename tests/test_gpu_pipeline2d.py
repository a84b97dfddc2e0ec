"""The reference's 2-D learned-closure pipeline (ai/pipeline.py:35-432) on the HIP kernels.

The data runs (Smagorinsky LES, and the finer BGK run block-averaged) and the validation run with a
trained closure are compared with the oracle's D2Q9 chain in float64: ``ref.equilibrium`` /
``ref.collide_bgk`` / ``ref.collide_smagorinsky_bgk`` / ``ref.stream`` / ``ref.macroscopic``, the
initial field restated in numpy from pipeline.py:35-65, and for the AI-LES collision torch's own
modules with the trained weights (the form inference.py:68-92 evaluates).  fp32 on the GPU against
fp64: velocities to 2e-6 absolute (|u| ≤ 0.1).  The end-to-end cases are the reference's own
(tests/test_ai_turbulence.py:266-312, 413-449).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref
from tensorlbm_amd import closure2d as c2
from tensorlbm_amd import pipeline2d as p2
from tests.test_gpu_closure2d import _strain_torch, _torch_model

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _init_field(nx, ny, mean_u=0.05):
    """pipeline.py:35-65 in float64: D2Q9 equilibrium at rho = 1."""
    yy, xx = np.meshgrid(np.arange(ny, dtype=np.float64), np.arange(nx, dtype=np.float64), indexing="ij")
    kx, ky = 2.0 * np.pi / nx, 2.0 * np.pi / ny
    ux = mean_u + 0.02 * (np.sin(2 * kx * xx) * np.cos(ky * yy) + 0.5 * np.sin(4 * kx * xx + 0.3) * np.cos(2 * ky * yy))
    uy = 0.02 * (np.cos(kx * xx) * np.sin(2 * ky * yy) + 0.5 * np.cos(3 * kx * xx) * np.sin(ky * yy + 0.7))
    return ref.equilibrium(ref.D2Q9, np.ones_like(ux), ux, uy)


def _block_mean(a, s):
    ny, nx = a.shape[0] // s, a.shape[1] // s
    return a[:ny * s, :nx * s].reshape(ny, s, nx, s).mean(axis=(1, 3))


def test_initial_field():
    got = p2._init_random_velocity_field(21, 13, 0, DEV).cpu().numpy()
    np.testing.assert_allclose(got, _init_field(21, 13), rtol=0, atol=2e-7)


def test_les_data_run_matches_oracle():
    snaps = p2._run_les_smoke(24, 20, 0.56, 0.17, 9, 4, 0, DEV)
    assert len(snaps) == 2
    f = _init_field(24, 20)
    want = []
    for step in range(9):
        f = ref.stream(ref.collide_smagorinsky_bgk(f, 0.56, 0.17))
        if (step + 1) % 4 == 0:
            want.append(ref.macroscopic(f)[1:3])
    for (gx, gy), (wx, wy) in zip(snaps, want):
        np.testing.assert_allclose(gx.cpu().numpy(), wx, rtol=0, atol=2e-6)
        np.testing.assert_allclose(gy.cpu().numpy(), wy, rtol=0, atol=2e-6)


@pytest.mark.parametrize("scale,shape", [(2, (20, 20)), (3, (11, 7))])
def test_dns_data_run_matches_oracle(scale, shape):
    nx, ny = shape
    snaps = p2._run_dns_reference(nx, ny, 0.6, 6, 3, 0, DEV, dns_scale=scale, warmup_steps=2)
    assert len(snaps) == 2 and tuple(snaps[0][0].shape) == (ny, nx)
    f = _init_field(nx * scale, ny * scale)
    for _ in range(2):
        f = ref.stream(ref.collide_bgk(f, 0.6))
    want = []
    for step in range(6):
        f = ref.stream(ref.collide_bgk(f, 0.6))
        if (step + 1) % 3 == 0:
            _, ux, uy = ref.macroscopic(f)[:3]
            want.append((_block_mean(ux, scale), _block_mean(uy, scale)))
    for (gx, gy), (wx, wy) in zip(snaps, want):
        np.testing.assert_allclose(gx.cpu().numpy(), wx, rtol=0, atol=2e-6)
        np.testing.assert_allclose(gy.cpu().numpy(), wy, rtol=0, atol=2e-6)


def test_coarsen_mean_drops_trailing_cells_and_checks_its_input():
    a = torch.arange(7 * 10, dtype=torch.float32, device=DEV).reshape(7, 10)
    got = p2._coarsen_mean(a, 3).cpu().numpy()
    np.testing.assert_allclose(got, _block_mean(a.cpu().numpy(), 3), rtol=1e-7)
    assert p2._coarsen_mean(a, 1) is a
    with pytest.raises(ValueError, match="too large"):
        p2._coarsen_mean(a, 8)
    with pytest.raises(ValueError, match="2-D"):
        p2._coarsen_mean(a.reshape(1, 7, 10), 2)


def test_validation_run_matches_oracle(tmp_path):
    torch.manual_seed(4)
    m = c2.EddyViscosityMLP(c2.ModelArch(3, 16, 2, "tanh"))
    m.set_feature_stats(torch.zeros(3), torch.full((3,), 2e-3))
    path = c2.save_model(m, tmp_path / "m.pt")
    got = p2._run_ai_validation(18, 14, 0.52, 5, 1, DEV, path)
    tm = _torch_model(m)
    f_ai = _init_field(18, 14)
    f_bgk = f_ai.copy()
    for _ in range(5):
        rho, vx, vy = ref.macroscopic(f_ai)[:3]
        sxx, syy, sxy = _strain_torch(torch.from_numpy(vx), torch.from_numpy(vy))
        nu = tm(torch.stack([sxx, syy, sxy], -1).reshape(-1, 3)).detach().reshape(14, 18).clamp_min(0.0).numpy()
        tau_eff = np.maximum(0.52 + 3.0 * nu, 0.501)
        feq = ref.equilibrium(ref.D2Q9, rho, vx, vy)
        f_ai = ref.stream(f_ai - (f_ai - feq) / tau_eff[None])
        f_bgk = ref.stream(ref.collide_bgk(f_bgk, 0.52))
    _, ux, uy = ref.macroscopic(f_ai)[:3]
    _, bx, by = ref.macroscopic(f_bgk)[:3]
    sx, sy, sz = _strain_torch(torch.from_numpy(ux), torch.from_numpy(uy))
    nu_t = tm(torch.stack([sx, sy, sz], -1).reshape(-1, 3)).detach().clamp_min(0.0).numpy()
    want = {"ai_umax": np.sqrt(ux * ux + uy * uy).max(), "bgk_umax": np.sqrt(bx * bx + by * by).max(),
            "ai_ke": (0.5 * (ux * ux + uy * uy)).mean(), "bgk_ke": (0.5 * (bx * bx + by * by)).mean(),
            "ai_nu_t_mean": nu_t.mean(), "ai_nu_t_max": nu_t.max()}
    assert got["stable"] is True
    assert 1e-4 < want["ai_nu_t_mean"] < 1.0  # the closure is active, so the two runs differ
    assert abs(want["ai_ke"] - want["bgk_ke"]) > 1e-3 * want["bgk_ke"]
    for k, v in want.items():
        assert got[k] == pytest.approx(float(v), rel=2e-5), k


def test_ai_les_is_stable_and_conserves_mass():
    """tests/test_ai_turbulence.py:235-284: a fresh model keeps the run finite and bounded, mass as BGK."""
    ux, uy = (t.to(DEV) for t in _sine(24, 24))
    f = tl.equilibrium(torch.ones_like(ux), ux, uy)
    model = c2.EddyViscosityMLP()
    f1, f2 = c2.collide_ai_les_bgk(f.clone(), 0.8, model), tl.collide_bgk(f.clone(), 0.8)
    assert torch.isfinite(f1).all() and torch.isfinite(f2).all()
    assert torch.allclose(tl.macroscopic(f1)[0].sum(), tl.macroscopic(f2)[0].sum(), rtol=1e-5)
    for _ in range(20):
        f = tl.stream(c2.collide_ai_les_bgk(f, 0.8, model))
    assert torch.isfinite(f).all()
    assert float(tl.macroscopic(f)[1].abs().max()) < 1.0


def _sine(nx, ny):
    yy, xx = torch.meshgrid(torch.arange(ny).float(), torch.arange(nx).float(), indexing="ij")
    kx, ky = 2.0 * torch.pi / nx, 2.0 * torch.pi / ny
    return (0.05 + 0.02 * torch.sin(kx * xx) * torch.cos(ky * yy), 0.02 * torch.cos(kx * xx) * torch.sin(ky * yy))


def _small_config():
    return c2.TrainConfig(epochs=3, batch_size=256, learning_rate=5e-3)


def test_pipeline_les_source(tmp_path):
    """tests/test_ai_turbulence.py:287-312 (the run database is not part of this package)."""
    res = tl.run_ai_les_pipeline(tmp_path, nx=20, ny=20, data_steps=8, sample_every=4, val_steps=4,
                                 train_config=_small_config(), seed=0)
    assert isinstance(res, tl.AIPipelineResult)
    assert res.dataset_path.exists() and res.model_path.exists()
    assert res.n_samples == 2 * 20 * 20 and res.n_snapshots == 2 and res.data_source == "les"
    assert res.validation["stable"] is True
    assert "history" in res.training and len(res.training["history"]) == 3
    assert res.db_path is None and res.to_dict()["model_path"] == str(res.model_path)
    ds = c2.load_dataset_pt(res.dataset_path)
    assert len(ds) == res.n_samples and ds.c_s == pytest.approx(0.1)
    assert c2.load_model(res.model_path).arch == c2.ModelArch()


def test_pipeline_dns_source_and_wrapper(tmp_path):
    """tests/test_ai_turbulence.py:413-449."""
    kw = dict(nx=20, ny=20, data_steps=8, sample_every=4, val_steps=4, dns_scale=2, dns_warmup_steps=2,
              train_config=_small_config(), seed=0)
    res = tl.run_ai_les_pipeline(tmp_path / "dns", data_source="dns", **kw)
    assert res.data_source == "dns" and res.n_snapshots == 2 and res.n_samples == 2 * 20 * 20
    assert res.validation["stable"] is True
    res2 = tl.run_ai_dns_pipeline(tmp_path / "dns_wrapper", **kw)
    assert res2.data_source == "dns" and res2.model_path.exists()
    # same seed, same data, same training: the wrapper is the same run
    assert [h["train_mse"] for h in res2.training["history"]] == [h["train_mse"] for h in res.training["history"]]
    with pytest.raises(ValueError, match="data_source"):
        tl.run_ai_les_pipeline(tmp_path / "bad", data_source="rans", **kw)
