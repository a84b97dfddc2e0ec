"""The learned-closure LES workload at full size (192³ LES trained on box-filtered 576³ DNS).

The 3-D form of the reference's ``run_ai_les_pipeline`` (ai/pipeline.py:222-398), end to end on
the GPU kernels: two 576³ D3Q19 BGK DNS runs from the random solenoidal field (k0 = 4,
u_rms = 0.03, τ = 0.51; one trains, one is the truth), 3³ block averages every 10 coarse steps,
strain features and eddy-viscosity labels, Adam training of the 7 → 32 → 32 → 1 closure, then
192³ LES runs of 300 steps from the filtered truth's first snapshot in fp16-deviation storage:
no closure, Smagorinsky C_s = 0.17 (turbulence.py:183-208) and the learned closure evaluated in
the fused step, with bf16 matrix-core operands (the configuration's "BF16 weights, FP32
accumulation") and with fp16 operands.  The learned-closure LES is repeated in fp64 storage.

Asserted: every run stays finite; the fp16-deviation closure run is within 1e-3 velocity rel-L2
of its fp64-storage twin; the learned closure's velocity error against the filtered DNS is below
Smagorinsky's, at the last snapshot and on average over the 30 snapshots.  At this spectrum the
192³ grid resolves the flow, and no closure at all does better than either model (every eddy
viscosity only adds dissipation there); the test prints all three.
"""

from __future__ import annotations

import math

import pytest
import torch

from tensorlbm_amd.les_pipeline import run_ai_les_pipeline_3d

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(600)
@pytest.mark.parametrize("operands", ["bf16", "fp16"])
def test_learned_closure_les_192_from_576_dns(dev, operands):
    res = run_ai_les_pipeline_3d(n_les=192, dns_scale=3, tau_dns=0.51, steps=300, sample_every=10, epochs=10,
                                 C_s=0.17, les_dtype=torch.float16, device=dev, k0=4.0, u_rms=0.03,
                                 max_samples=1 << 24, precision_check=True, log=lambda m: print(m, flush=True),
                                 closure_operands=operands)
    les = res["les"]
    for name, r in les.items():
        print(f"{name:12s} velocity rel-L2 final {r['err_final']:.4e} mean {r['err_mean']:.4e} "
              f"KE error final {r['ke_err_final']:+.4e}")
        assert all(math.isfinite(v) for v in r.values()), name
    print(f"learned closure fp16-deviation vs fp64 storage: {res['learned_vs_fp64_rel_l2']:.3e}; "
          f"val R2 {res['val_r2']:.3f}, {res['samples_train']} training samples")
    assert res["n_dns"] == 576 and res["samples_train"] > 1_000_000
    assert res["learned_vs_fp64_rel_l2"] <= 1e-3
    assert les["learned"]["err_final"] < les["smagorinsky"]["err_final"]
    assert les["learned"]["err_mean"] < les["smagorinsky"]["err_mean"]
