"""The HIP ops on the reference's own input/output vectors (``tests/golden/reference_vectors.npz``).

The vectors were written by the reference's functions (``tests/golden/make_reference_vectors.py``);
the same inputs go through the C ABI here and must reproduce them:

* ``collide_cumulant_d3q27`` (legacy pseudo-cumulant operator, cumulant.py:935-1047), fp64,
  with and without the domain-mean Smagorinsky rate;
* ``collide_mrt27`` with every rate distinct (d3q27.py:643-726), fp32 as the reference computes it;
* ``collide_cumulant_d3q19`` with the WALE / Vreman closures, solid masks and wall velocities
  (cumulant.py:1159-1377), fp64 and fp32.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref

pytestmark = pytest.mark.gpu

VECTORS = np.load(Path(__file__).parent / "golden" / "reference_vectors.npz")


def _gpu(a, dev, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return t if dtype is None else t.to(dtype)


@pytest.mark.parametrize("i", range(5))
def test_legacy_d3q27_reproduces_the_reference(i, dev):
    f, p, want = VECTORS[f"legacy27_{i}_in"], VECTORS[f"legacy27_{i}_params"], VECTORS[f"legacy27_{i}_out"]
    tau, ob, oo, oe, cs = (float(v) for v in p)
    got = tl.collide_cumulant_d3q27(_gpu(f, dev), tau, omega_b=ob, omega_odd=oo, omega_even=oe, C_s=cs)
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=2e-14)


@pytest.mark.parametrize("i", range(3))
def test_mrt27_reproduces_the_reference(i, dev):
    f, p, want = VECTORS[f"mrt27_{i}_in"], VECTORS[f"mrt27_{i}_params"], VECTORS[f"mrt27_{i}_out"]
    s_pi = None if np.isnan(p[4]) else float(p[4])
    got = tl.collide_mrt27(_gpu(f, dev), float(p[0]), s_e=float(p[1]), s_eps=float(p[2]), s_q=float(p[3]), s_pi=s_pi)
    assert got.dtype == torch.float32
    # fp32: the reference applies fp32-rounded matrices, the kernel exact rational tables
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=5e-7)


def _sgs(i, dev, dtype):
    p = VECTORS[f"sgs19_{i}_params"]
    mask = None
    if bool(VECTORS[f"sgs19_{i}_has_mask"]):
        mask = torch.from_numpy(VECTORS[f"sgs19_{i}_mask"]).to(dev)
    kw = dict(omega_b=float(p[1]), omega_even=float(p[2]), C_w=float(p[3]), C_v=float(p[4]), solid_mask=mask,
              wall_velocity=tuple(float(v) for v in p[5:8]))
    return _gpu(VECTORS[f"sgs19_{i}_in"], dev, dtype), float(p[0]), kw


@pytest.mark.parametrize("i", range(6))
def test_wale_vreman_cumulant_d3q19_reproduces_the_reference_fp64(i, dev):
    f, tau, kw = _sgs(i, dev, torch.float64)
    got = tl.collide_cumulant_d3q19(f, tau, **kw)
    np.testing.assert_allclose(got.cpu().numpy(), VECTORS[f"sgs19_{i}_out"], rtol=0, atol=2e-14)


@pytest.mark.parametrize("i", range(6))
def test_wale_vreman_cumulant_d3q19_fp32(i, dev):
    f, tau, kw = _sgs(i, dev, torch.float32)
    got = tl.collide_cumulant_d3q19(f, tau, **kw)
    assert got.dtype == torch.float32
    np.testing.assert_allclose(got.cpu().numpy(), VECTORS[f"sgs19_{i}_out"], rtol=0, atol=1e-6)


@pytest.mark.parametrize("model", ["wale", "vreman"])
def test_gradient_closure_tau_matches_the_oracle(model, dev):
    from tests import _fields as F

    f = F.random_state(19, (6, 7, 9), seed=12, noise=2e-3, u_amp=0.05)
    mask = np.zeros((6, 7, 9), bool)
    mask[2:4, 3:5, 1:6] = True
    kw = {"C_w": 0.5} if model == "wale" else {"C_v": 0.07}
    want = ref.gradient_closure_tau(f, 0.53, solid_mask=mask, wall_velocity=(0.01, 0.0, -0.02), **kw)
    got = tl.gradient_closure_tau(_gpu(f, dev), 0.53, solid_mask=torch.from_numpy(mask).to(dev),
                                  wall_velocity=(0.01, 0.0, -0.02), **kw)
    assert got.dtype == torch.float64 and tuple(got.shape) == (6, 7, 9)
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=1e-14)
    assert np.ptp(want) > 1e-3  # the closure varies over the grid


def test_gradient_closure_feeds_the_fused_step(dev):
    """The fused obstacle step with the closure's tau_eff field (float32, TLBM_TAU_FIELD) equals the
    oracle chain stream -> cumulant collision at that per-cell rate."""
    from tests import _fields as F

    n = (8, 9, 10)
    tau = 0.55
    f0 = F.random_state(19, n, seed=3, noise=2e-3, u_amp=0.05).astype(np.float32)
    fs = ref.stream(f0.astype(np.float64))
    te = ref.gradient_closure_tau(fs, tau, C_w=0.5)
    want = ref.collide_cumulant_d3q19(fs, tau, omega=1.0 / te.astype(np.float32).astype(np.float64))
    te_gpu = tl.gradient_closure_tau(tl.stream3d(_gpu(f0, dev)), tau, C_w=0.5)
    assert te_gpu.dtype == torch.float32
    np.testing.assert_allclose(te_gpu.cpu().numpy(), te, rtol=2e-6)
    got = tl.fused_obstacle_step(_gpu(f0, dev), (tau - 0.5) / 3.0, torch.zeros(n, dtype=torch.bool, device=dev),
                                 0.0, 1.0, collision="CUMULANT", tau_eff=te_gpu)
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=2e-6)


def test_sgs_argument_errors(dev):
    f = _gpu(VECTORS["sgs19_0_in"], dev)
    with pytest.raises(ValueError, match="only one SGS model"):
        tl.collide_cumulant_d3q19(f, 0.6, C_w=0.5, C_v=0.02)
    with pytest.raises(ValueError, match="non-negative"):
        tl.collide_cumulant_d3q19(f, 0.6, C_w=-0.5)
    with pytest.raises(ValueError, match="wall_velocity"):
        tl.collide_cumulant_d3q19(f, 0.6, C_w=0.5, wall_velocity=(0.0, float("nan"), 0.0))
    with pytest.raises(ValueError, match="solid_mask"):
        tl.collide_cumulant_d3q19(f, 0.6, C_w=0.5, solid_mask=torch.zeros(f.shape[1:], dtype=torch.int8, device=dev))
