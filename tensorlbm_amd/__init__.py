"""tensorlbm_amd — TensorLBM's per-step lattice update, native on MI355X (gfx950).

The public names mirror the reference package ``tensorlbm`` for the hot path
(equilibrium / macroscopic / collide_* / stream*, bounce-back, Smagorinsky,
obstacle force, the fused step and its solver object).  Every op runs a
hand-written HIP kernel from ``libtlbm_hip.so`` through its C ABI
(``include/tlbm_hip.h``); there is no CPU or eager fallback.
"""

from __future__ import annotations

import torch  # noqa: F401  (loads torch's HIP runtime before the library binds to it)

from . import lattice
from .fused import (
    FusedSolver3D,
    apply_far_field_bc_6face,
    apply_mass_correction,
    fused_obstacle_step,
    fused_step,
    hip_collide,
    hip_fused,
    hip_stream,
    is_available,
)
from .cavities import BatchedCavities
from .checkpoint import (
    CheckpointError,
    SolverCheckpoint,
    absolute_populations,
    eager_load_state_dict,
    eager_state_dict,
    fused_load_state_dict,
    fused_state_dict,
    load_checkpoint,
    load_solver_checkpoint,
    save_checkpoint,
    save_solver_checkpoint,
    triton_fused_load_state_dict,
    triton_fused_state_dict,
)
from .closure import EddyViscosityMLP3D, TrainConfig3D, train_eddy_viscosity_model_3d
from .closure2d import (
    EddyViscosityMLP,
    collide_ai_les_bgk,
    extract_les_samples_2d,
    predict_nu_t_2d,
    predict_tau_eff_2d,
    strain_rate_tensor_2d,
    train_eddy_viscosity_model,
)
from .diagnostics import frontal_area, obstacle_drag_coefficient, voxel_wetted_area
from .executor import LBMStepExecutor
from .ops import *  # noqa: F401,F403
from .ops import __all__ as _ops_all
from .pipeline2d import AIPipelineResult, run_ai_dns_pipeline, run_ai_les_pipeline
from .precision import PrecisionPolicy, cast_to_compute, cast_to_store
from .reporters import CallbackReporter, StepContext, ThroughputReporter

# reference module names for the fused GPU path (src/tensorlbm/triton_fused.py,
# src/tensorlbm/triton_fused_obstacle.py); here they run the HIP kernels
triton_fused = hip_fused
triton_stream = hip_stream
triton_collide = hip_collide
triton_fused_obstacle_xfar_les = fused_obstacle_step
TritonFusedSolver3D = FusedSolver3D

# the reference's top-level velocity-set tables (src/tensorlbm/__init__.py:265-272): D2Q9 C/W/OPPOSITE,
# D3Q19 C3D/W3D/OPPOSITE3D, D3Q27 C27/W27/OPPOSITE27; ``lattice.D3Q19.W_EXACT64`` etc. hold the fp64 weights
C, W, OPPOSITE = lattice.D2Q9.C, lattice.D2Q9.W, lattice.D2Q9.OPPOSITE
C3D, W3D, OPPOSITE3D = lattice.D3Q19.C, lattice.D3Q19.W, lattice.D3Q19.OPPOSITE
C27, W27, OPPOSITE27 = lattice.D3Q27.C, lattice.D3Q27.W, lattice.D3Q27.OPPOSITE

__version__ = "0.1.0"

__all__ = list(_ops_all) + [
    "C", "W", "OPPOSITE", "C3D", "W3D", "OPPOSITE3D", "C27", "W27", "OPPOSITE27",
    "EddyViscosityMLP", "collide_ai_les_bgk", "extract_les_samples_2d", "predict_nu_t_2d", "predict_tau_eff_2d",
    "strain_rate_tensor_2d", "train_eddy_viscosity_model", "AIPipelineResult", "run_ai_les_pipeline",
    "run_ai_dns_pipeline", "lattice", "FusedSolver3D", "fused_obstacle_step", "apply_far_field_bc_6face", "apply_mass_correction", "fused_step", "hip_collide", "hip_fused", "hip_stream",
    "is_available", "PrecisionPolicy", "cast_to_compute", "cast_to_store", "StepContext", "ThroughputReporter",
    "CallbackReporter", "triton_fused", "triton_stream", "triton_collide", "triton_fused_obstacle_xfar_les",
    "TritonFusedSolver3D", "LBMStepExecutor", "BatchedCavities", "EddyViscosityMLP3D", "TrainConfig3D",
    "train_eddy_viscosity_model_3d", "CheckpointError", "SolverCheckpoint", "absolute_populations",
    "eager_load_state_dict", "eager_state_dict", "fused_load_state_dict", "fused_state_dict", "load_checkpoint",
    "load_solver_checkpoint", "save_checkpoint", "save_solver_checkpoint", "triton_fused_load_state_dict",
    "triton_fused_state_dict", "frontal_area", "obstacle_drag_coefficient", "voxel_wetted_area",
]
