"""The eager-torch CPU restatement used as the CPU baseline equals the numpy oracle's chain."""

from __future__ import annotations

import numpy as np
import torch

from oracle import ref, torch_cpu
from tests import _fields as F


def test_channel_step_equals_the_oracle_chain():
    n = (6, 7, 9)
    f0 = F.random_state(19, n, seed=5, noise=2e-3)
    solid = ref.sphere_mask(9, 7, 6, 4.0, 3.0, 2.5, 1.8)
    tau, u_in = 0.57, 0.05
    want = f0
    got = torch.from_numpy(f0)
    obstacle = torch.from_numpy(solid)
    for _ in range(3):
        want = ref.channel_bc(ref.stream(ref.collide_bgk(want, tau)), u_in, solid=solid)
        got = torch_cpu.channel_step(got, tau, u_in, obstacle)
    assert solid.any()
    np.testing.assert_allclose(got.numpy(), want, rtol=0, atol=1e-14)
