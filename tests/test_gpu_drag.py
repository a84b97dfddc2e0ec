"""Drag coefficient of a sphere in the fused channel: force sampled in-kernel, normalised on the host.

From equilibrium everywhere (solid cells included, as the reference's cases
start), 100 fused steps of the channel + sphere give the oracle's momentum-
exchange force (fp64 oracle, fp32 kernel: 1e-3 relative); the sphere is
symmetric about the channel axis, so the lateral components stay at rounding
level.  C_D from the GPU mask equals C_D from the same mask on the CPU (the area
count is exact integer work on either device).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref

pytestmark = pytest.mark.gpu


def _oracle_force(f, solid, u_in, tau, steps):
    force = None
    for _ in range(steps):
        fin = ref.stream(f)
        force = ref.obstacle_force(fin, solid)
        f = ref.collide_bgk(ref.channel_bc(fin, u_in, solid), tau)
    return force


def test_sphere_channel_drag_coefficient():
    nz, ny, nx, u_in, tau, steps = 20, 20, 64, 0.05, 0.6, 100
    solid = ref.sphere_mask(nx, ny, nz, 20.0, 10.0, 10.0, 4.5)
    mask = torch.from_numpy(solid).to(torch.int8)
    one = np.ones((nz, ny, nx))
    f0 = ref.equilibrium(ref.D3Q19, one, u_in * one, 0.0 * one, 0.0 * one)
    s = tl.FusedSolver3D(nz, ny, nx, tau, "cuda:0", solid=mask, boundary="channel_x", u_bc=(u_in, 0.0, 0.0))
    f = torch.from_numpy(f0).to("cuda:0", torch.float32)
    force = torch.zeros(3, dtype=torch.float64, device="cuda:0")
    for _ in range(steps):
        force.zero_()
        f = s.step(f, force=force)
    fx, fy, fz = force.tolist()
    want = _oracle_force(f0, solid, u_in, tau, steps)
    assert want[0] > 0.1
    assert abs(fx - want[0]) <= 1e-3 * want[0]
    assert abs(fy) <= 1e-3 * want[0] and abs(fz) <= 1e-3 * want[0]
    gpu = tl.obstacle_drag_coefficient(fx, s.solid, u_in)
    cpu = tl.obstacle_drag_coefficient(fx, mask, u_in)
    assert gpu == cpu and gpu["s_ref_kind"] == "wetted" and gpu["C_D"] > 0.0
    assert gpu["C_D"] == pytest.approx(fx / (0.5 * u_in * u_in * tl.voxel_wetted_area(mask)))
    assert tl.obstacle_drag_coefficient(fx, s.solid, u_in, s_ref="frontal")["S_ref"] == tl.frontal_area(mask)
