"""Host-side checks of the production companions and the D3Q27 ops (no GPU, no compute): argument
validation happens before any launch, and CPU tensors are refused (the HIP path has no CPU fallback)."""

from __future__ import annotations

import pytest
import torch

import tensorlbm_amd as tl


def test_far_field_6face_checks_shape_then_device():
    with pytest.raises(ValueError, match="Q >= 19"):
        tl.apply_far_field_bc_6face(torch.zeros(9, 3, 3, 3), 0.05)
    with pytest.raises(ValueError, match="Q >= 19"):
        tl.apply_far_field_bc_6face(torch.zeros(19, 3, 3), 0.05)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        tl.apply_far_field_bc_6face(torch.zeros(19, 3, 3, 3), 0.05)


def test_mass_correction_checks_shape_then_device():
    with pytest.raises(ValueError, match="19 or 27"):
        tl.apply_mass_correction(torch.zeros(9, 3, 3, 3), 1.0)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        tl.apply_mass_correction(torch.ones(27, 2, 2, 2), 1.0)


def test_d3q27_ops_refuse_wrong_lattice_and_cpu():
    with pytest.raises(ValueError, match="27"):
        tl.compute_obstacle_forces_27(torch.zeros(19, 2, 2, 2), torch.zeros(2, 2, 2, dtype=torch.bool))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        tl.compute_obstacle_forces_27(torch.zeros(27, 2, 2, 2), torch.zeros(2, 2, 2, dtype=torch.bool))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        tl.collide_smagorinsky_mrt27(torch.zeros(27, 2, 2, 2), 0.7)


def test_fused_obstacle_step_refuses_cpu():
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        tl.fused_obstacle_step(torch.zeros(19, 2, 2, 2), 0.01, torch.zeros(2, 2, 2, dtype=torch.int8), 0.1, 1.0)


def test_reference_table_names():
    assert tl.C.shape == (9, 2) and tl.C3D.shape == (19, 3) and tl.C27.shape == (27, 3)
    assert torch.equal(tl.OPPOSITE3D, tl.lattice.D3Q19.OPPOSITE) and torch.equal(tl.W27, tl.lattice.D3Q27.W)
    for name in ("C", "W", "OPPOSITE", "C3D", "W3D", "OPPOSITE3D", "C27", "W27", "OPPOSITE27",
                 "apply_far_field_bc_6face", "apply_mass_correction", "collide_smagorinsky_mrt27",
                 "compute_obstacle_forces_27"):
        assert name in tl.__all__, name
