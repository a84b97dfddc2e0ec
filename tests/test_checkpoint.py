"""Checkpoint files and state dicts (host logic; no GPU).

Mirrors the reference's checkpoint tests (tests/test_checkpoint.py:44-436): the
tensor + JSON pair, the tagged solver envelope with its integrity digests, and
the fail-closed adapters.  Solvers are attribute stubs with the fields the
adapters read, as in the reference's own adapter tests.  Beyond the reference:
the deviation-storage offset convention (``storage_offset``) and the check that
a state written the reference's way (no offset key) restores into a plain solver.
"""

from __future__ import annotations

import json
from types import SimpleNamespace

import pytest
import torch

import tensorlbm_amd as tl
from tensorlbm_amd import checkpoint as ck
from tensorlbm_amd.lattice import D3Q19


def _solver(**kw):
    base = dict(lattice="D3Q19", Q=19, nz=4, ny=5, nx=6, tau=0.6, dtype=torch.float32, device=torch.device("cpu"),
                precision_policy=None, deviation=False, collision="BGK", boundary=None, _report_step=0)
    base.update(kw)
    return SimpleNamespace(**base)


def _f(dtype=torch.float32, shape=(19, 4, 5, 6), seed=0):
    g = torch.Generator().manual_seed(seed)
    return (1.0 / 19 + 1e-3 * torch.randn(shape, generator=g)).to(dtype)


# ----------------------------------------------------------------- tensor + JSON pair


class TestPair:
    def test_round_trip(self, tmp_path):
        f = _f()
        assert tl.save_checkpoint(f, 42, tmp_path / "a" / "b", extra={"re": 100.0, "label": "x"}) == tmp_path / "a" / "b"
        g, step, meta = tl.load_checkpoint(tmp_path / "a" / "b")
        assert torch.equal(g, f) and g.device.type == "cpu" and step == 42
        assert meta["re"] == 100.0 and meta["label"] == "x" and meta["step"] == 42

    def test_files_and_keys_are_the_reference_ones(self, tmp_path):
        tl.save_checkpoint(_f(torch.float16), 7, tmp_path)
        assert sorted(p.name for p in tmp_path.iterdir()) == ["checkpoint_f.pt", "checkpoint_meta.json"]
        meta = json.loads((tmp_path / "checkpoint_meta.json").read_text())
        assert meta == {"format_version": 1, "step": 7, "tensor_shape": [19, 4, 5, 6], "tensor_dtype": "torch.float16",
                        "lattice_directions": 19}

    def test_view_saves_only_its_elements(self, tmp_path):
        big = torch.arange(19 * 4 * 70, dtype=torch.float32).reshape(19, 4, 70)
        view = big[:, :, 3:9]  # strided, as the solver's padded buffers are
        tl.save_checkpoint(view, 1, tmp_path)
        g, _, _ = tl.load_checkpoint(tmp_path)
        assert torch.equal(g, view) and g.untyped_storage().nbytes() == view.numel() * 4

    @pytest.mark.parametrize("name", ["checkpoint_f.pt", "checkpoint_meta.json"])
    def test_missing_file(self, tmp_path, name):
        tl.save_checkpoint(_f(), 1, tmp_path)
        (tmp_path / name).unlink()
        with pytest.raises(FileNotFoundError):
            tl.load_checkpoint(tmp_path)

    @pytest.mark.parametrize("meta, match", [
        ("{not json", "valid JSON"),
        ("[1, 2]", "JSON object"),
        ('{"format_version": 1}', "missing 'step'"),
        ('{"format_version": 1, "step": 3.5}', "'step' must be an integer"),
        ('{"format_version": 1, "step": true}', "'step' must be an integer"),
        ('{"format_version": 2, "step": 3}', "Unsupported checkpoint format_version=2"),
        ('{"format_version": "1", "step": 3}', "'format_version' must be an integer"),
        ('{"format_version": 1, "step": 3, "tensor_shape": [19, 4, 5, 7]}', "does not match metadata"),
    ])
    def test_bad_metadata(self, tmp_path, meta, match):
        tl.save_checkpoint(_f(), 1, tmp_path)
        (tmp_path / "checkpoint_meta.json").write_text(meta)
        with pytest.raises(ValueError, match=match):
            tl.load_checkpoint(tmp_path)

    def test_version_zero_is_accepted(self, tmp_path):
        tl.save_checkpoint(_f(), 1, tmp_path)
        (tmp_path / "checkpoint_meta.json").write_text('{"step": 5}')
        assert tl.load_checkpoint(tmp_path)[1] == 5

    def test_expected_shape_and_directions(self, tmp_path):
        tl.save_checkpoint(_f(), 1, tmp_path)
        tl.load_checkpoint(tmp_path, expected_shape=(19, 4, 5, 6), expected_lattice_directions=19)
        with pytest.raises(ValueError, match="current run shape"):
            tl.load_checkpoint(tmp_path, expected_shape=(19, 4, 5, 7))
        with pytest.raises(ValueError, match="expected 27 directions, got 19"):
            tl.load_checkpoint(tmp_path, expected_lattice_directions=27)

    def test_non_tensor_payload(self, tmp_path):
        tl.save_checkpoint(_f(), 1, tmp_path)
        torch.save({"f": 1}, tmp_path / "checkpoint_f.pt")
        with pytest.raises(ValueError, match="not a torch.Tensor"):
            tl.load_checkpoint(tmp_path)

    def test_no_temporary_files(self, tmp_path):
        for step in range(3):
            tl.save_checkpoint(_f(seed=step), step, tmp_path)
        assert sorted(p.name for p in tmp_path.iterdir()) == ["checkpoint_f.pt", "checkpoint_meta.json"]
        assert tl.load_checkpoint(tmp_path)[1] == 2


# ----------------------------------------------------------------- solver envelope


class TestEnvelope:
    def _save(self, tmp_path, state=None, **kw):
        state = state if state is not None else tl.eager_state_dict(_solver(), _f(), step=3)
        return tl.save_solver_checkpoint(tmp_path / "s.ckpt", state, **kw)

    def test_round_trip(self, tmp_path):
        solver, f = _solver(), _f()
        mask = torch.zeros(4, 5, 6, dtype=torch.int8)
        mask[1:3, 2, 2:4] = 1
        solver.mask = mask
        path = self._save(tmp_path, tl.eager_state_dict(solver, f, step=9, extra={"case": "box"}),
                          metadata={"note": "n"})
        c = tl.load_solver_checkpoint(path)
        assert c.step == 9 and c.lattice == "D3Q19" and c.grid == (4, 5, 6) and c.q == 19
        assert c.format_version == 2 and c.metadata == {"note": "n"} and c.path == path
        assert torch.equal(c.populations, f) and torch.equal(c.state["obstacle_mask"], mask)
        assert c.state["extra"] == {"case": "box"} and c.state["tau"] == 0.6
        other = _solver(_report_step=0)
        g = tl.eager_load_state_dict(other, c.state)
        assert torch.equal(g, f) and other._report_step == 9
        assert sorted(p.name for p in tmp_path.iterdir()) == ["s.ckpt"]

    def test_step_defaults_to_the_run_counter(self):
        assert tl.eager_state_dict(_solver(_report_step=17), _f())["step"] == 17

    def test_bf16_digest(self, tmp_path):
        f = _f(torch.bfloat16)
        path = self._save(tmp_path, tl.fused_state_dict(_solver(dtype=torch.bfloat16), f, step=1))
        assert torch.equal(tl.load_solver_checkpoint(path).populations, f)

    def test_digest_mismatch(self, tmp_path):
        path = self._save(tmp_path)
        env = torch.load(path, weights_only=True)
        env["state"]["populations"][0, 0, 0, 0] += 1.0
        torch.save(env, path)
        with pytest.raises(ck.CheckpointError, match="integrity digest for 'populations'"):
            tl.load_solver_checkpoint(path)

    def test_invalid_digest_record(self, tmp_path):
        path = self._save(tmp_path)
        env = torch.load(path, weights_only=True)
        env["digests"]["obstacle_mask"] = "00"
        torch.save(env, path)
        with pytest.raises(ck.CheckpointError, match="invalid integrity record"):
            tl.load_solver_checkpoint(path)

    def test_truncated_and_garbage(self, tmp_path):
        path = self._save(tmp_path)
        data = path.read_bytes()
        path.write_bytes(data[: len(data) // 2])
        with pytest.raises(ck.CheckpointError, match="could not be decoded"):
            tl.load_solver_checkpoint(path)
        path.write_bytes(b"not a checkpoint")
        with pytest.raises(ck.CheckpointError, match="could not be decoded"):
            tl.load_solver_checkpoint(path)
        with pytest.raises(ck.CheckpointError, match="not found"):
            tl.load_solver_checkpoint(tmp_path / "absent.ckpt")

    @pytest.mark.parametrize("edit, match", [
        (lambda e: e.update(format_version=3), "format_version=3"),
        (lambda e: e.pop("format"), "missing format tag"),
        (lambda e: e["state"].pop("step"), "missing required key"),
    ])
    def test_envelope_fields(self, tmp_path, edit, match):
        path = self._save(tmp_path)
        env = torch.load(path, weights_only=True)
        edit(env)
        torch.save(env, path)
        with pytest.raises(ck.CheckpointError, match=match):
            tl.load_solver_checkpoint(path)

    def test_expected_identity(self, tmp_path):
        path = self._save(tmp_path)
        tl.load_solver_checkpoint(path, expected_lattice="D3Q19", expected_grid=(4, 5, 6), expected_q=19)
        for kw, match in (({"expected_lattice": "D3Q27"}, "lattice"), ({"expected_grid": (4, 5, 7)}, "grid"),
                          ({"expected_q": 27}, "Q mismatch")):
            with pytest.raises(ck.CheckpointError, match=match):
                tl.load_solver_checkpoint(path, **kw)

    def test_rng_restored(self, tmp_path):
        torch.manual_seed(1234)
        path = self._save(tmp_path)
        want = torch.rand(5)
        torch.manual_seed(99)
        tl.load_solver_checkpoint(path)
        assert torch.equal(torch.rand(5), want)
        torch.manual_seed(99)
        tl.load_solver_checkpoint(path, restore_rng=False)
        assert not torch.equal(torch.rand(5), want)

    @pytest.mark.parametrize("state, match", [
        ({"step": 1}, "missing required key"),
        ({"populations": [1.0], "step": 1}, "must be a torch.Tensor"),
        ({"populations": torch.tensor(1.0), "step": 1}, "lattice dimension"),
        ({"populations": torch.ones(19), "step": -1}, "non-negative integer"),
        ({"populations": torch.ones(19), "step": True}, "non-negative integer"),
        ({"populations": torch.ones(19), "step": 1, "grid": (4, 0, 6)}, "positive ints"),
        ({"populations": torch.ones(19), "step": 1, "lattice": 19}, "string or None"),
        ({"populations": torch.ones(19), "step": 1, "q": 0}, "positive integer"),
        ("populations", "must be a dict"),
    ])
    def test_state_validation(self, tmp_path, state, match):
        with pytest.raises(ck.CheckpointError, match=match):
            tl.save_solver_checkpoint(tmp_path / "x.ckpt", state)


# ----------------------------------------------------------------- fused adapters


class TestFusedAdapters:
    def test_records_identity(self):
        policy = tl.PrecisionPolicy.FP32FP16
        s = _solver(precision_policy=policy, dtype=torch.float16, deviation=True, collision="MRT",
                    boundary="channel_x")
        st = tl.fused_state_dict(s, _f(torch.float16), step=4)
        assert st["lattice"] == "D3Q19" and st["q"] == 19 and st["grid"] == (4, 5, 6) and st["step"] == 4
        assert st["precision_policy"] == "FP32FP16" and st["storage_dtype"] == "torch.float16:f-w"
        assert st["dtype"] == "torch.float16" and st["storage_offset"] == "f-w"
        assert st["collision"] == "MRT" and st["boundary"] == "channel_x"
        assert tl.fused_state_dict(_solver(), _f())["storage_offset"] == "none"
        assert tl.triton_fused_state_dict is tl.fused_state_dict

    def test_lattice_defaults_to_d3q19(self):
        assert tl.fused_state_dict(_solver(lattice=None), _f())["lattice"] == "D3Q19"

    def test_d2q9_grid(self):
        s = _solver(lattice="D2Q9", Q=9, nz=1)
        st = tl.fused_state_dict(s, _f(shape=(9, 5, 6)))
        assert st["grid"] == (5, 6)
        assert tl.fused_load_state_dict(_solver(lattice="D2Q9", Q=9, nz=1), st).shape == (9, 5, 6)

    def test_round_trip_restores_step(self):
        f = _f(torch.float16)
        s = _solver(dtype=torch.float16, deviation=True)
        other = _solver(dtype=torch.float16, deviation=True)
        g = tl.fused_load_state_dict(other, tl.fused_state_dict(s, f, step=11))
        assert torch.equal(g, f) and other._report_step == 11

    def test_reference_style_state_loads_into_plain_solver(self):
        # the keys the reference's triton_fused_state_dict writes (checkpoint.py:580-585): no offset key
        f = _f()
        state = {"populations": f, "step": 2, "lattice": "D3Q19", "grid": (4, 5, 6), "q": 19, "tau": 0.6,
                 "dtype": "torch.float32", "obstacle_mask": None, "precision_policy": None,
                 "storage_dtype": "torch.float32"}
        assert torch.equal(tl.fused_load_state_dict(_solver(), state), f)
        with pytest.raises(ck.CheckpointError, match="storage offset 'none'"):
            tl.fused_load_state_dict(_solver(deviation=True), state)

    def test_deviation_state_refused_by_plain_solver(self):
        st = tl.fused_state_dict(_solver(dtype=torch.float16, deviation=True), _f(torch.float16))
        with pytest.raises(ck.CheckpointError, match="storage offset 'f-w'"):
            tl.fused_load_state_dict(_solver(dtype=torch.float16), st)

    def test_policy_mismatch(self):
        st = tl.fused_state_dict(_solver(precision_policy=tl.PrecisionPolicy.FP32FP16, dtype=torch.float16),
                                 _f(torch.float16))
        with pytest.raises(ck.CheckpointError, match="precision policy"):
            tl.fused_load_state_dict(_solver(dtype=torch.float16), st)

    def test_dtype_mismatch(self):
        st = tl.fused_state_dict(_solver(dtype=torch.float16), _f(torch.float16))
        with pytest.raises(ck.CheckpointError, match="storage dtype"):
            tl.fused_load_state_dict(_solver(), st)
        st = tl.fused_state_dict(_solver(), _f())
        st["populations"] = st["populations"].double()
        with pytest.raises(ck.CheckpointError, match="record says"):
            tl.fused_load_state_dict(_solver(), st)

    @pytest.mark.parametrize("solver_kw, match", [
        ({"lattice": "D3Q27", "Q": 27}, "lattice"),
        ({"lattice": None, "Q": 27}, "Q 19"),
        ({"nx": 7}, "grid"),
        ({"tau": 0.61}, "tau"),
        ({"collision": "CUMULANT"}, "collision"),
        ({"boundary": "lid_y"}, "boundary"),
    ])
    def test_identity_mismatch(self, solver_kw, match):
        st = tl.fused_state_dict(_solver(boundary="channel_x"), _f())
        with pytest.raises(ck.CheckpointError, match=match):
            tl.fused_load_state_dict(_solver(**{"boundary": "channel_x", **solver_kw}), st)

    def test_populations_shape_checked_against_grid(self):
        st = tl.fused_state_dict(_solver(), _f())
        st["populations"] = _f(shape=(19, 4, 5, 7))
        with pytest.raises(ck.CheckpointError, match="grid"):
            tl.fused_load_state_dict(_solver(), st)

    def test_absolute_populations(self):
        f = _f(torch.float16, seed=5)
        st = tl.fused_state_dict(_solver(dtype=torch.float16, deviation=True), f)
        want = f.float() + D3Q19.W.reshape(19, 1, 1, 1)
        assert torch.equal(tl.absolute_populations(st), want)
        plain = tl.fused_state_dict(_solver(), _f())
        assert torch.equal(tl.absolute_populations(plain), plain["populations"])
        st["storage_offset"] = "f-feq"
        with pytest.raises(ck.CheckpointError, match="unknown storage offset"):
            tl.absolute_populations(st)

    def test_envelope_keeps_the_offset(self, tmp_path):
        st = tl.fused_state_dict(_solver(dtype=torch.bfloat16, deviation=True), _f(torch.bfloat16), step=6)
        c = tl.load_solver_checkpoint(tl.save_solver_checkpoint(tmp_path / "d.ckpt", st))
        assert c.state["storage_offset"] == "f-w" and c.state["storage_dtype"] == "torch.bfloat16:f-w"
        g = tl.fused_load_state_dict(_solver(dtype=torch.bfloat16, deviation=True), c.state)
        assert torch.equal(g, st["populations"])

    def test_reference_loader_refuses_a_deviation_state(self):
        """The reference's fused loader (checkpoint.py:613-642) checks only the precision tier and the
        storage dtype string.  A deviation state must fail the second check there, so that f - w is
        never read as f; a plain fp16 state from the same tier must pass both."""

        def reference_checks(solver, state):
            if state.get("precision_policy") != getattr(getattr(solver, "precision_policy", None), "name", None):
                raise ck.CheckpointError("precision policy")
            saved = state.get("storage_dtype") or state.get("dtype")
            if str(saved) != str(solver.dtype):
                raise ck.CheckpointError("storage dtype")

        ref_solver = _solver(precision_policy=tl.PrecisionPolicy.FP32FP16, dtype=torch.float16)
        dev = tl.fused_state_dict(_solver(precision_policy=tl.PrecisionPolicy.FP32FP16, dtype=torch.float16,
                                          deviation=True), _f(torch.float16))
        with pytest.raises(ck.CheckpointError, match="storage dtype"):
            reference_checks(ref_solver, dev)
        plain = tl.fused_state_dict(ref_solver, _f(torch.float16))
        reference_checks(ref_solver, plain)
        assert plain["storage_dtype"] == "torch.float16" and plain["storage_offset"] == "none"

    def test_storage_dtype_and_offset_must_agree(self):
        st = tl.fused_state_dict(_solver(dtype=torch.float16, deviation=True), _f(torch.float16))
        st["storage_offset"] = "none"
        with pytest.raises(ck.CheckpointError, match="names offset 'f-w'"):
            tl.fused_load_state_dict(_solver(dtype=torch.float16, deviation=True), st)
