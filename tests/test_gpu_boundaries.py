"""Boundary conditions as separate ops (the eager forms of the fused boundaries) against the oracle.

``apply_simple_channel_boundaries_3d`` (boundaries3d.py:407-443), ``far_field_bc_3d`` (:446-537) and
``zou_he_moving_lid_3d`` (:324-404) run one HIP kernel over the post-stream state; ``sphere_mask``
(:29-49) builds the obstacle on the GPU; ``collide_bgk_fused`` / ``collide_bgk_matmul`` (solver.py:74-246)
are the BGK update of ``collide_bgk``.  fp64 to 1e-13, fp32 to 1e-6; masks bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref
from tests import _fields as F

pytestmark = pytest.mark.gpu

TOL = {torch.float64: 1e-13, torch.float32: 1e-6}


def _state(shape, seed, dtype):
    return torch.from_numpy(F.random_state(19, shape, seed=seed, noise=3e-3, u_amp=0.03)).to("cuda:0", dtype)


@pytest.mark.parametrize("radius,centre,shape", [
    (5.05, (5.05, 4.0, 4.0), (12, 12, 14)), (7.7, (5.0, 4.3, 4.0), (12, 12, 14)), (3.3, (5.2, 4.5, 4.0), (12, 12, 14)),
    (24.0, (64.0, 128.0, 128.0), (256, 256, 256))])
def test_sphere_mask_bit_exact(radius, centre, shape):
    nz, ny, nx = shape
    got = tl.sphere_mask(nx, ny, nz, *centre, radius, "cuda:0")
    assert got.dtype == torch.bool and tuple(got.shape) == shape
    np.testing.assert_array_equal(got.cpu().numpy(), ref.sphere_mask(nx, ny, nz, *centre, radius))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_channel_boundaries(dtype):
    shape = (9, 10, 17)
    f = _state(shape, 31, dtype)
    rng = np.random.default_rng(5)
    wall = np.zeros(shape, dtype=bool)
    wall[:, 0, :] = wall[:, -1, :] = True
    obstacle = ref.sphere_mask(17, 10, 9, 8.0, 5.0, 4.0, 2.5) | (rng.random(shape) < 0.05)
    assert (wall & obstacle).any()  # cells in both masks swap twice
    f0 = f.clone()
    got = tl.apply_simple_channel_boundaries_3d(f, 0.04, torch.from_numpy(wall), torch.from_numpy(obstacle))
    assert torch.equal(f, f0)  # input untouched
    want = ref.channel_bc(f.cpu().numpy(), 0.04, solid=obstacle, wall=wall)
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=TOL[dtype])


@pytest.mark.parametrize("config", [None, {"far_field_faces": ["y-", "y+"], "periodic_faces": ["z-", "z+"]},
                                    {"far_field_faces": ["z-", "y+"], "periodic_faces": ["x-", "x+"]},
                                    {"far_field_faces": [], "periodic_faces": ["x+"]}])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_far_field(config, dtype):
    shape = (8, 9, 13)
    f = _state(shape, 32, dtype)
    obstacle = ref.sphere_mask(13, 9, 8, 6.0, 4.0, 4.0, 2.2)
    got = tl.far_field_bc_3d(f, 0.05, torch.from_numpy(obstacle), uy=0.01, uz=-0.005, bc_config=config)
    want = ref.far_field_bc(f.cpu().numpy(), 0.05, obstacle, uy=0.01, uz=-0.005, bc_config=config)
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=TOL[dtype])
    got_free = tl.far_field_bc_3d(f, 0.05, bc_config=config)
    want_free = ref.far_field_bc(f.cpu().numpy(), 0.05, None, bc_config=config)
    np.testing.assert_allclose(got_free.cpu().numpy(), want_free, rtol=0, atol=TOL[dtype])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_zou_he_lid_every_cell_of_the_plane(dtype):
    shape = (7, 11, 9)
    f = _state(shape, 33, dtype)
    got = tl.zou_he_moving_lid_3d(f, 0.08)
    want = ref.zou_he_moving_lid(f.cpu().numpy(), 0.08)
    np.testing.assert_allclose(got.cpu().numpy(), want, rtol=0, atol=TOL[dtype])
    # the lid plane's x-momentum is rho u_lid on every cell, corners included
    rho, ux, _, _ = ref.macroscopic(got.double().cpu().numpy())
    np.testing.assert_allclose((rho * ux)[:, -1, :], 0.08 * rho[:, -1, :], rtol=0, atol=10 * TOL[dtype])
    with pytest.raises(NotImplementedError):
        tl.zou_he_moving_lid_3d(f, 0.08, uy_lid=0.01)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_bgk_fused_and_matmul_forms(dtype):
    f = torch.from_numpy(F.random_state(9, (12, 16), seed=34)).to("cuda:0", dtype)
    base = tl.collide_bgk(f, 0.63)
    assert torch.equal(tl.collide_bgk_fused(f, 0.63), base)
    assert torch.equal(tl.collide_bgk_matmul(f, 0.63), base)
    np.testing.assert_allclose(base.cpu().numpy(), ref.collide_bgk(f.cpu().numpy(), 0.63), rtol=0, atol=TOL[dtype])


def test_argument_errors():
    f = _state((4, 5, 2), 35, torch.float32)
    with pytest.raises(ValueError):
        tl.apply_simple_channel_boundaries_3d(f, 0.04, torch.zeros((4, 5, 2), dtype=torch.bool),
                                              torch.zeros((4, 5, 2), dtype=torch.bool))  # nx < 3
    g = _state((4, 5, 6), 36, torch.float32)
    with pytest.raises(ValueError, match="mask shape"):
        tl.far_field_bc_3d(g, 0.05, torch.zeros((4, 5, 5), dtype=torch.bool))
    with pytest.raises(ValueError, match="Unsupported dtype"):
        tl.zou_he_moving_lid_3d(g.half(), 0.05)
