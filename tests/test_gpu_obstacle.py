"""GPU parity of the fused obstacle step (full-way bounce-back, ω dispatch, Ladd force).

Mirrors tests/test_triton_obstacle_wall_parity.py of the reference: single-step
parity < 1e-6 for BGK and CM (here also CUMULANT), the integer-encoded gather
canary, and the force-buffer parity.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref
from tests import _fields as F

pytestmark = pytest.mark.gpu

NX, NY, NZ = 32, 16, 16
U_IN, NU = 0.05, 0.005
TAU = 3.0 * NU + 0.5


def _case(noise=0.0):
    solid = ref.sphere_mask(NX, NY, NZ, 12.0, 8.0, 8.0, 4.5)
    rho = np.ones((NZ, NY, NX))
    ux = np.full_like(rho, U_IN)
    ux[solid] = 0.0
    f0 = ref.equilibrium(ref.D3Q19, rho, ux, np.zeros_like(rho), np.zeros_like(rho))
    if noise:
        rng = np.random.default_rng(4)
        f0 = f0 + noise * rng.standard_normal(f0.shape) * ref.D3Q19.w.reshape(-1, 1, 1, 1)
    return solid, f0.astype(np.float32)


def _gpu(a, dev, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device=dev, dtype=dtype)


@pytest.mark.parametrize("collision", ["BGK", "CM", "CUMULANT"])
def test_single_step_parity_full_way_bb(collision, dev):
    solid, f0 = _case(noise=2e-2)
    out = tl.fused_obstacle_step(_gpu(f0, dev), NU, _gpu(solid, dev, torch.int8), 0.0, 1.0, collision=collision)
    want, _ = ref.fused_obstacle_step(f0.astype(np.float64), TAU, solid, collision=collision)
    d = np.abs(out.cpu().numpy() - want)
    assert d.max() < 1e-6, d.max()
    assert d[:, solid].max() < 1e-6


def test_gather_addressing_bitwise_canary(dev):
    solid, _ = _case()
    fc = F.canary(19, NZ, NY, NX).astype(np.float32)
    out = tl.fused_obstacle_step(_gpu(fc, dev), 1e9, _gpu(solid, dev, torch.int8), 0.0, 1.0)
    got = out.cpu().numpy().round().astype(np.int64)
    want = ref.fullway_pull(fc.astype(np.float64), solid).round().astype(np.int64)
    assert int((got != want).sum()) == 0


def test_force_buffer_matches_ladd(dev):
    solid, f0 = _case(noise=1e-2)
    fx, fy, fz = (torch.zeros((), device=dev) for _ in range(3))
    tl.fused_obstacle_step(_gpu(f0, dev), NU, _gpu(solid, dev, torch.int8), 0.0, 1.0, fx_buf=fx, fy_buf=fy,
                           fz_buf=fz)
    # production: force on the streamed state, sampled before bounce-back
    want = ref.obstacle_force(ref.stream(f0.astype(np.float64)), solid)
    scale = abs(want[0])
    for got, w, lim in ((fx, want[0], 1e-5), (fy, want[1], 1e-4), (fz, want[2], 1e-4)):
        assert abs(float(got) - w) < lim * scale + 1e-7, (float(got), w)


def test_force_accumulates_in_fp64_buffer(dev):
    solid, f0 = _case(noise=1e-2)
    acc = torch.zeros(3, dtype=torch.float64, device=dev)
    fg = _gpu(f0, dev)
    tl.fused_obstacle_step(fg, NU, _gpu(solid, dev, torch.int8), 0.0, 1.0, force=acc)
    tl.fused_obstacle_step(fg, NU, _gpu(solid, dev, torch.int8), 0.0, 1.0, force=acc)
    want = ref.obstacle_force(ref.stream(f0.astype(np.float64)), solid)
    assert np.allclose(acc.cpu().numpy(), 2.0 * np.array(want), rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("collision", ["BGK", "CM", "CUMULANT"])
def test_internal_smagorinsky_matches_oracle(collision, dev):
    solid, f0 = _case(noise=5e-2)
    cs = 0.17
    out = tl.fused_obstacle_step(_gpu(f0, dev), NU, _gpu(solid, dev, torch.int8), cs, 1.0, collision=collision)
    want, _ = ref.fused_obstacle_step(f0.astype(np.float64), TAU, solid, collision=collision, cs_delta_sq=cs * cs)
    plain, _ = ref.fused_obstacle_step(f0.astype(np.float64), TAU, solid, collision=collision)
    assert np.abs(want - plain).max() > 1e-5  # the closure is not a no-op here
    assert np.abs(out.cpu().numpy() - want).max() < 1e-6


def test_external_tau_eff(dev):
    solid, f0 = _case(noise=2e-2)
    rng = np.random.default_rng(3)
    te = (0.52 + 0.3 * rng.random((NZ, NY, NX))).astype(np.float32)
    out = tl.fused_obstacle_step(_gpu(f0, dev), NU, _gpu(solid, dev, torch.int8), 0.0, 1.0, tau_eff=_gpu(te, dev))
    want, _ = ref.fused_obstacle_step(f0.astype(np.float64), TAU, solid, tau_eff=te.astype(np.float64))
    assert np.abs(out.cpu().numpy() - want).max() < 1e-6


def test_channel_sphere_multistep_vs_production_chain(dev):
    """Fused channel step vs the eager chain collide -> stream -> inlet/outlet -> bounce-back."""
    nx, ny, nz, tau, u_in = 40, 20, 18, 0.56, 0.05
    solid = ref.sphere_mask(nx, ny, nz, 12.0, 10.0, 9.0, 4.0)
    rng = np.random.default_rng(0)
    rho = np.ones((nz, ny, nx))
    ux = u_in + 1e-3 * rng.uniform(-1, 1, rho.shape)
    ux[solid] = 0.0
    f0 = ref.equilibrium(ref.D3Q19, rho, ux, 1e-3 * rng.uniform(-1, 1, rho.shape),
                         1e-3 * rng.uniform(-1, 1, rho.shape))
    # the fused step stores b = C(h); start from b0 = C(h0)
    h = f0
    b_ref = ref.collide_bgk(h, tau)
    solver = tl.FusedSolver3D(nz, ny, nx, tau=tau, device=dev, dtype=torch.float64, solid=_gpu(solid, dev, torch.int8),
                              boundary="channel_x", u_bc=(u_in, 0.0, 0.0))
    b = _gpu(b_ref, dev, torch.float64)
    for _ in range(15):
        h = ref.channel_bc(ref.stream(ref.collide_bgk(h, tau)), u_in, solid=solid)
        b = solver.step(b)
    want = ref.collide_bgk(h, tau)
    got = b.cpu().numpy()
    assert np.allclose(got, want, rtol=1e-11, atol=1e-13), np.abs(got - want).max()


@pytest.mark.parametrize("bc_config", [None, {"far_field_faces": ["y-", "z+"], "periodic_faces": ["x+"]},
                                       {"far_field_faces": ["y+", "z-"], "periodic_faces": ["x-", "z-"]}])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_fused_far_field_matches_oracle_chain(dev, bc_config, dtype):
    """stream → far_field_bc_3d (boundaries3d.py:446-537) → bounce-back → BGK, in one launch."""
    solid, f0 = _case(noise=1e-2)
    tau, uin = 0.56, (0.05, 0.01, -0.02)
    s = tl.FusedSolver3D(NZ, NY, NX, tau, dev, dtype, solid=torch.tensor(solid, device=dev), boundary="far_field",
                         u_bc=uin, bc_config=bc_config, layout="contiguous")
    f = _gpu(f0, dev, dtype)
    want = f0.astype(np.float64)
    for _ in range(3):
        f = s.step(f)
        fin = ref.far_field_bc(ref.stream(want), uin[0], solid, uin[1], uin[2], bc_config)
        feq, _, _ = ref._feq_safe(fin)
        want = fin - (fin - feq) / tau
    tol = dict(rtol=1e-5, atol=1e-7) if dtype == torch.float32 else dict(rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(f.cpu().numpy(), want, **tol)


# The one-step kernel's inlet (x = 0) and outlet (x = nx - 1) lanes take the pulled state of x = 1 /
# x = nx - 2 from the lane of their wave that gathered it.  nx - 1 a multiple of 64 puts the outlet's
# neighbour in the previous wave (loaded instead); lanes past nx (ragged 64/128/256-wide tiles) act
# as outlet lanes; x-periodic far-field faces keep their own state.
EDGE_NX = [33, 65, 70, 129, 130, 200]
FAR_CFG = {"channel_x": None, "far_field": {"far_field_faces": ["y-", "y+"], "periodic_faces": ["z-", "z+"]},
           "far_field_xper": {"far_field_faces": ["y-", "z+"], "periodic_faces": ["x+"]},
           "far_field_xper2": {"far_field_faces": ["y+"], "periodic_faces": ["x-", "z-", "z+"]}}


@pytest.mark.parametrize("nx", EDGE_NX)
@pytest.mark.parametrize("boundary", list(FAR_CFG))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_one_step_inlet_outlet_lanes(dev, nx, boundary, dtype):
    ny, nz, tau = 3, 4, 0.56
    uin = (0.05, 0.01, -0.02)
    solid = ref.sphere_mask(nx, ny, nz, nx / 3.0, 1.0, 2.0, 1.3)
    f0 = F.random_state(19, (nz, ny, nx), seed=nx, noise=2e-2, u_amp=0.05)
    far = boundary.startswith("far")
    s = tl.FusedSolver3D(nz, ny, nx, tau, dev, dtype, solid=torch.tensor(solid, device=dev),
                         boundary="far_field" if far else "channel_x", u_bc=uin if far else (uin[0], 0.0, 0.0),
                         bc_config=FAR_CFG[boundary], layout="contiguous", temporal_blocking=False)
    got = s.step(_gpu(f0, dev, dtype)).cpu().numpy()
    f_in = f0.astype(np.float32).astype(np.float64) if dtype == torch.float32 else f0  # what the GPU was given
    fin = ref.stream(f_in)
    if far:
        fb = ref.far_field_bc(fin, uin[0], solid, uin[1], uin[2], FAR_CFG[boundary])
    else:
        fb = ref.channel_bc(fin, uin[0], solid=solid)
    feq, _, _ = ref._feq_safe(fb)
    want = fb - (fb - feq) / tau
    tol = dict(rtol=1e-5, atol=1e-7) if dtype == torch.float32 else dict(rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(got, want, **tol)
    # the edge planes are where a wrong lane would show
    np.testing.assert_allclose(got[..., 0], want[..., 0], **tol)
    np.testing.assert_allclose(got[..., -1], want[..., -1], **tol)
