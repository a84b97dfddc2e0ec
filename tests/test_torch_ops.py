"""torch.ops.tlbm registers with torch's dispatcher and traces without a GPU (no compute here)."""

from __future__ import annotations

import pytest
import torch

from tensorlbm_amd import _native as nat
from tensorlbm_amd import torch_ops


def test_schemas_registered():
    t = torch_ops.load()
    for name in ("equilibrium", "macroscopic", "stream", "collide", "step"):
        assert hasattr(t, name), name
    schema = str(t.step.default._schema)
    assert "Tensor(a!)? force=None" in schema and "int steps=1" in schema


def test_meta_shapes():
    t = torch_ops.load()
    for q, shape in ((9, (5, 7)), (19, (3, 5, 7)), (27, (3, 5, 7))):
        f = torch.empty((q, *shape), device="meta")
        assert t.stream(f).shape == f.shape
        assert t.collide(f, nat.COLL_BGK, 0.6).shape == f.shape
        assert t.step(f, nat.COLL_BGK, 0.6, steps=2).shape == f.shape
        parts = t.macroscopic(f)
        assert len(parts) == (3 if q == 9 else 4) and all(p.shape == shape for p in parts)
        fields = [torch.empty(shape, device="meta") for _ in range(3 if q == 9 else 4)]
        uz = fields[3] if q != 9 else None
        assert t.equilibrium(q, fields[0], fields[1], fields[2], uz).shape == f.shape


def test_cpu_tensors_raise_runtime_error():
    t = torch_ops.load()
    with pytest.raises(RuntimeError, match="GPU device"):
        t.step(torch.zeros((19, 2, 2, 2)), nat.COLL_BGK, 0.6)
    with pytest.raises(RuntimeError, match="GPU device"):
        t.equilibrium(9, torch.ones(2, 2), torch.zeros(2, 2), torch.zeros(2, 2))
