"""Drag normalisation: wetted / frontal voxel areas and C_D (host logic, CPU tensors).

The reference's cases (tests/test_drag_survey.py:72-76, tests/test_hydrodynamics.py:24-30):
a 4³ block has wetted area 96 and frontal area 16; one voxel at dx = 0.5 has 1.5,
two joined voxels 2.5.  The face count is pinned further by brute force over
random masks (6 neighbours per solid cell), by the exact identity of an
orthogonally convex body (every axis-parallel line through it crosses its
surface twice, so wetted = 2 (A_x + A_y + A_z)), and by the voxel sphere's
frontal area tending to πr².
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref


def _brute_faces(m: np.ndarray) -> int:
    nz, ny, nx = m.shape
    n = 0
    for z, y, x in zip(*np.nonzero(m)):
        for dz, dy, dx in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            a, b, c = z + dz, y + dy, x + dx
            if not (0 <= a < nz and 0 <= b < ny and 0 <= c < nx) or not m[a, b, c]:
                n += 1
    return n


def test_reference_block_and_voxels():
    mask = torch.zeros((10, 10, 12), dtype=torch.bool)
    mask[3:7, 3:7, 4:8] = True
    assert tl.voxel_wetted_area(mask) == 96 and tl.frontal_area(mask) == 16
    one = torch.zeros((3, 3, 3), dtype=torch.bool)
    one[1, 1, 1] = True
    assert tl.voxel_wetted_area(one, 0.5) == pytest.approx(1.5)
    one[1, 1, 2] = True  # touches the domain's x face: that face counts too
    assert tl.voxel_wetted_area(one, 0.5) == pytest.approx(2.5)


@pytest.mark.parametrize("seed", range(6))
def test_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    m = rng.random((5, 6, 7)) < (0.2 + 0.12 * seed)
    assert tl.voxel_wetted_area(torch.from_numpy(m)) == _brute_faces(m)
    assert tl.voxel_wetted_area(torch.from_numpy(m.astype(np.int8)), 2.0) == 4 * _brute_faces(m)


def test_full_domain_counts_the_outside_faces():
    m = torch.ones((3, 4, 5), dtype=torch.int8)
    assert tl.voxel_wetted_area(m) == 2 * (4 * 5 + 3 * 5 + 3 * 4)


@pytest.mark.parametrize("axes", [(7.5, 7.5, 7.5), (11.2, 6.3, 4.1), (3.0, 9.5, 5.2)])
def test_orthoconvex_identity(axes):
    nz, ny, nx = 30, 26, 34
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    rx, ry, rz = axes
    m = torch.from_numpy(((x - 15.3) / rx) ** 2 + ((y - 12.9) / ry) ** 2 + ((z - 14.1) / rz) ** 2 <= 1.0)
    projections = sum(tl.frontal_area(m, a) for a in range(3))
    assert tl.voxel_wetted_area(m) == 2 * projections


def test_sphere_frontal_area_tends_to_pi_r2():
    r, n = 24.7, 56
    m = torch.from_numpy(ref.sphere_mask(n, n, n, n / 2, n / 2, n / 2, r))
    assert abs(tl.frontal_area(m) / (math.pi * r * r) - 1.0) < 1e-2
    assert abs(tl.voxel_wetted_area(m) / (6.0 * math.pi * r * r) - 1.0) < 1e-2


def test_drag_coefficient():
    mask = torch.zeros((10, 10, 12), dtype=torch.bool)
    mask[3:7, 3:7, 4:8] = True
    d = tl.obstacle_drag_coefficient(0.12, mask, 0.05)
    q = 0.5 * 1.0 * 0.05 ** 2 * 96
    assert d == {"C_D": pytest.approx(0.12 / q), "S_ref": 96.0, "s_ref_kind": "wetted",
                 "dynamic_pressure": pytest.approx(q), "fx": 0.12}
    f = tl.obstacle_drag_coefficient(0.12, mask, 0.05, s_ref="frontal", rho_lu=1.02, dx=0.5)
    assert f["S_ref"] == 4.0 and f["s_ref_kind"] == "frontal"
    assert f["C_D"] == pytest.approx(0.12 / (0.5 * 1.02 * 0.05 ** 2 * 4.0))
    e = tl.obstacle_drag_coefficient(0.12, mask, 0.05, s_ref_value=10.0, s_ref="ignored")
    assert e["s_ref_kind"] == "explicit" and e["S_ref"] == 10.0
    assert tl.obstacle_drag_coefficient(0.12, mask, 0.0)["C_D"] == 0.0
    with pytest.raises(ValueError, match="s_ref"):
        tl.obstacle_drag_coefficient(0.12, mask, 0.05, s_ref="side")
    with pytest.raises(ValueError, match="3D"):
        tl.voxel_wetted_area(mask[0])
    with pytest.raises(ValueError, match="streamwise_axis"):
        tl.frontal_area(mask, 3)
