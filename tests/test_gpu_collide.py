"""GPU parity of the eager (one-launch) ops against the oracle, fp32 and fp64."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref
from tests import _fields as F

pytestmark = pytest.mark.gpu

SHAPES = {9: (1, 20, 26), 19: (6, 10, 14), 27: (5, 8, 12)}
TOL = {torch.float32: dict(rtol=1e-5, atol=2e-7), torch.float64: dict(rtol=1e-12, atol=1e-14)}


def _state(q, seed=1, noise=1e-2):
    f = F.random_state(q, SHAPES[q], seed=seed, noise=noise)
    return f[:, 0] if q == 9 else f


def _gpu(a, dev, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dtype)


def _cmp(got: torch.Tensor, want: np.ndarray, dtype):
    g = got.cpu().numpy()
    assert np.allclose(g, want, **TOL[dtype]), float(np.abs(g - want).max())


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("q", [9, 19, 27])
def test_equilibrium_and_macroscopic(q, dtype, dev):
    rng = np.random.default_rng(q)
    sh = SHAPES[q][1:] if q == 9 else SHAPES[q]
    rho = 1.0 + 0.05 * rng.standard_normal(sh)
    u = 0.05 * rng.standard_normal((3,) + sh)
    lat = F.LATS[q]
    if q == 9:
        got = tl.equilibrium(*(_gpu(a, dev, dtype) for a in (rho, u[0], u[1])))
        want = ref.equilibrium(lat, rho, u[0], u[1])
    else:
        fn = tl.equilibrium3d if q == 19 else tl.equilibrium27
        got = fn(*(_gpu(a, dev, dtype) for a in (rho, u[0], u[1], u[2])))
        want = ref.equilibrium(lat, rho, u[0], u[1], u[2])
    _cmp(got, want, dtype)
    f = _state(q)
    mfn = {9: tl.macroscopic, 19: tl.macroscopic3d, 27: tl.macroscopic27}[q]
    outs = mfn(_gpu(f, dev, dtype))
    wants = ref.macroscopic(f)
    for g, w in zip(outs, wants):
        _cmp(g, w, dtype)


def test_equilibrium_shape_mismatch_raises(dev):
    a = torch.ones(4, 4, 4, device=dev)
    with pytest.raises(ValueError, match="shapes must match"):
        tl.equilibrium3d(a, a, a, torch.ones(4, 4, 3, device=dev))


def test_equilibrium_out_argument(dev):
    a = torch.ones(3, 4, 5, device=dev)
    out = torch.empty(19, 3, 4, 5, device=dev)
    res = tl.equilibrium3d(a, 0.01 * a, 0.0 * a, 0.0 * a, out=out)
    assert res.data_ptr() == out.data_ptr()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("q", [9, 19, 27])
def test_bgk(q, dtype, dev):
    f = _state(q)
    fn = {9: tl.collide_bgk, 19: tl.collide_bgk3d, 27: tl.collide_bgk27}[q]
    _cmp(fn(_gpu(f, dev, dtype), 0.63), ref.collide_bgk(f, 0.63), dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("q", [9, 19, 27])
def test_mrt(q, dtype, dev):
    f = _state(q)
    fn = {9: tl.collide_mrt, 19: tl.collide_mrt3d, 27: tl.collide_mrt27}[q]
    _cmp(fn(_gpu(f, dev, dtype), 0.58), ref.collide_mrt(f, 0.58), dtype)


def test_mrt_d2q9_tau_field(dev):
    f = _state(9)
    rng = np.random.default_rng(5)
    tf = 0.55 + 0.3 * rng.random(f.shape[1:])
    got = tl.collide_mrt(_gpu(f, dev, torch.float64), 0.6, tau_field=_gpu(tf, dev, torch.float64))
    # cell-wise shear rate = moments 7, 8 relaxed at 1/tau_field
    want = np.empty_like(f)
    for j in range(f.shape[1]):
        for i in range(f.shape[2]):
            want[:, j, i] = ref.collide_mrt(f[:, j : j + 1, i : i + 1], float(np.float32(tf[j, i])))[:, 0, 0]
    assert np.allclose(got.cpu().numpy(), want, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_cumulant_d2q9(dtype, dev):
    f = _state(9, noise=3e-2)
    got = tl.collide_cumulant_d2q9(_gpu(f, dev, dtype), 0.61, omega_b=1.3, omega_3=0.7, omega_4=1.6)
    _cmp(got, ref.collide_cumulant_d2q9(f, 0.61, 1.3, 0.7, 1.6), dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_cascaded_d3q19(dtype, dev):
    f = _state(19)
    got = tl.collide_cascaded_d3q19(_gpu(f, dev, dtype), 0.57, s_bulk=1.3, s_3=1.1, s_4=0.9)
    _cmp(got, ref.collide_cascaded(f, 0.57, 1.3, 1.1, 0.9), dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_cascaded_d3q27(dtype, dev):
    f = _state(27)
    got = tl.collide_cascaded_d3q27(_gpu(f, dev, dtype), 0.57, s_bulk=1.3, s_3=1.1, s_4=0.9, s_5=1.2, s_6=1.4)
    _cmp(got, ref.collide_cascaded(f, 0.57, 1.3, 1.1, 0.9, 1.2, 1.4), dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("cs", [0.0, 0.15])
def test_cumulant_d3q19(cs, dtype, dev):
    f = _state(19, noise=5e-2)
    got = tl.collide_cumulant_d3q19(_gpu(f, dev, dtype), 0.56, omega_b=1.2, omega_even=0.8, C_s=cs)
    _cmp(got, ref.collide_cumulant_d3q19(f, 0.56, 1.2, 1.0, 0.8, C_s=cs), dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("cs", [0.0, 0.12])
def test_cumulant_d3q27_legacy(cs, dtype, dev):
    f = _state(27, noise=3e-2)
    got = tl.collide_cumulant_d3q27(_gpu(f, dev, dtype), 0.83, omega_b=1.13, omega_odd=0.71, omega_even=1.37, C_s=cs)
    want = ref.collide_cumulant_d3q27(f, 0.83, 1.13, 0.71, 1.37, C_s=cs)
    _cmp(got, want, dtype)


@pytest.mark.parametrize("q", [19, 27])
def test_smagorinsky_tau_sum(q, dev):
    from tensorlbm_amd import _native as nat

    f = _state(q, noise=5e-2)
    acc = torch.zeros(1, dtype=torch.float64, device=dev)
    fg = _gpu(f, dev, torch.float64)
    nat.check(nat.lib().tlbm_smagorinsky_tau_sum(q, nat.STORE_F64, fg.data_ptr(), None, *f.shape[1:], 0.6, 0.09,
                                                 acc.data_ptr(), nat.stream_of(fg)), "tau sum")
    want = ref.smagorinsky_tau_mean(f, 0.6, 0.3) * f[0].size
    assert abs(float(acc) - want) < 1e-10 * want


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("q", [9, 19])
def test_smagorinsky_mrt(q, dtype, dev):
    f = _state(q, noise=5e-2)
    if q == 9:
        got = tl.collide_smagorinsky_mrt(_gpu(f, dev, dtype), 0.58, C_s=0.2)
    else:
        got = tl.collide_smagorinsky_mrt3d(_gpu(f, dev, dtype), 0.58, C_s=0.2, s_pi=1.3)
    _cmp(got, ref.collide_smagorinsky_mrt(f, 0.58, 0.2, s_pi=None if q == 9 else 1.3), dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("cs", [0.0, 0.2])
def test_geier_d3q27(dtype, cs, dev):
    f = _state(27, noise=3e-2)
    got = tl.collide_cumulant_geier_d3q27(_gpu(f, dev, dtype), 0.58, omega_b=1.1, omega_odd=1.3, omega_even=0.9,
                                          C_s=cs)
    want = ref.collide_cumulant_geier_d3q27_les(f, 0.58, 1.1, 1.3, 0.9, C_s=cs)
    tol = dict(rtol=1e-4, atol=1e-6) if dtype == torch.float32 else TOL[dtype]
    g = got.cpu().numpy()
    assert np.allclose(g, want, **tol), float(np.abs(g - want).max())


@pytest.mark.parametrize("q", [9, 19, 27])
def test_smagorinsky_bgk(q, dev):
    f = _state(q, noise=8e-2)
    fn = {9: tl.collide_smagorinsky_bgk, 19: tl.collide_smagorinsky_bgk3d, 27: tl.collide_smagorinsky_bgk27}[q]
    got = fn(_gpu(f, dev, torch.float64), 0.53, C_s=0.17)
    want = ref.collide_smagorinsky_bgk(f, 0.53, 0.17)
    assert np.abs(want - ref.collide_bgk(f, 0.53)).max() > 1e-6
    assert np.allclose(got.cpu().numpy(), want, rtol=1e-11, atol=1e-13)


@pytest.mark.parametrize("q", [9, 19, 27])
def test_bounce_back(q, dev):
    f = _state(q)
    rng = np.random.default_rng(2)
    mask = rng.random(f.shape[1:]) < 0.3
    fn = {9: tl.bounce_back_cells, 19: tl.bounce_back_cells_3d, 27: tl.bounce_back_cells_27}[q]
    got = fn(_gpu(f, dev, torch.float64), _gpu(mask, dev, torch.bool))
    assert np.array_equal(got.cpu().numpy(), ref.bounce_back(f, mask))


def test_obstacle_force_op(dev):
    f = _state(19)
    rng = np.random.default_rng(8)
    mask = rng.random(f.shape[1:]) < 0.2
    fx, fy, fz = tl.compute_obstacle_forces_3d(_gpu(f, dev, torch.float64), _gpu(mask, dev, torch.bool))
    want = ref.obstacle_force(f, mask)
    assert np.allclose([float(fx), float(fy), float(fz)], want, rtol=1e-12, atol=1e-14)


def test_ops_reject_cpu_tensors():
    with pytest.raises(RuntimeError, match="GPU"):
        tl.collide_bgk3d(torch.ones(19, 2, 2, 2), 0.6)
