"""Host-side parts of the 2-D closure and its pipeline: file formats, the split, safe loading and the
refusal of CPU devices (tests/test_ai_turbulence.py:102-125, 379-410)."""

from __future__ import annotations

import json
import pickle

import pytest
import torch

from tensorlbm_amd import closure2d as c2
from tensorlbm_amd import pipeline2d as p2


def test_dataset_save_load_roundtrip(tmp_path):
    feats, target = torch.randn(50, 3), torch.rand(50, 1)
    ds = c2.EddyViscosityDataset(feats, target, c_s=0.12, description="unit-test")
    p = c2.save_dataset_pt(ds, tmp_path / "ds.pt")
    meta = json.loads(p.with_suffix(".pt.json").read_text())
    assert meta["format_version"] == 2 and meta["c_s"] == pytest.approx(0.12)
    back = c2.load_dataset_pt(p)
    assert torch.equal(back.features, feats) and torch.equal(back.targets, target)
    assert back.c_s == pytest.approx(0.12) and "unit-test" in back.description


def test_dataset_split():
    ds = c2.EddyViscosityDataset(torch.randn(100, 3), torch.randn(100, 1))
    tr, va = ds.split(0.2, seed=42)
    assert len(tr) + len(va) == 100 and len(va) == 20
    tr2, va2 = ds.split(0.2, seed=42)
    assert torch.equal(va.features, va2.features)  # seeded
    together = torch.cat([tr.features, va.features])
    assert torch.equal(together[together[:, 0].argsort()], ds.features[ds.features[:, 0].argsort()])
    with pytest.raises(ValueError):
        ds.split(1.0)


class _Evil:
    def __init__(self, target):
        self.target = target

    def __reduce__(self):
        return (exec, (f"from pathlib import Path; Path({str(self.target)!r}).write_text('boom')",))


def test_model_loader_refuses_pickled_code(tmp_path):
    target = tmp_path / "model-side-effect.txt"
    payload = tmp_path / "bad-model.pt"
    torch.save(_Evil(target), payload)
    payload.with_suffix(".pt.json").write_text(json.dumps({"arch": {"in_features": 3}, "format_version": 2}))
    with pytest.raises((pickle.UnpicklingError, RuntimeError, ValueError)):
        c2.load_model(payload)
    assert not target.exists()


def test_dataset_loader_refuses_pickled_code(tmp_path):
    target = tmp_path / "dataset-side-effect.txt"
    payload = tmp_path / "bad-dataset.pt"
    torch.save(_Evil(target), payload)
    payload.with_suffix(".pt.json").write_text(json.dumps({"c_s": 0.1, "description": "bad", "format_version": 2}))
    with pytest.raises((pickle.UnpicklingError, RuntimeError, ValueError)):
        c2.load_dataset_pt(payload)
    assert not target.exists()


def test_pipeline_refuses_a_cpu_device(tmp_path):
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        p2.run_ai_les_pipeline(tmp_path, device="cpu")
    with pytest.raises(ValueError, match="data_source"):
        p2.run_ai_les_pipeline(tmp_path, data_source="rans")


def test_multi_snapshot_samples_of_nothing_are_empty():
    f, t = c2.extract_les_samples_2d_multi([])
    assert tuple(f.shape) == (0, 3) and tuple(t.shape) == (0, 1)
