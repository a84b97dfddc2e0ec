"""Batched lid-driven cavities: one launch (or one graph) for many cavities.

Per cavity the batch must equal the single-grid fused step with the wall mask and the
moving-lid boundary (bit for bit in fp32) and the oracle's cavity chain (fp64).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref
from tensorlbm_amd.cavities import BatchedCavities, cavity_batch_config, wall_mask
from tests import _fields as F

pytestmark = pytest.mark.gpu

EDGES = [8, 11, 13, 16, 5]
TAUS = [0.8, 0.55, 0.62, 0.56, 1.3]
U_LID = 0.1


def _batch(dev, dtype=torch.float64, deviation=False, seed=0):
    b = BatchedCavities(EDGES, TAUS, u_lid=U_LID, device=dev, dtype=dtype, deviation=deviation)
    states = []
    for i, e in enumerate(EDGES):
        f0 = F.random_state(19, (e, e, e), seed=seed + i, noise=5e-3, u_amp=0.03)
        states.append(f0)
        w = ref.D3Q19.w.reshape(19, 1, 1, 1) if deviation else 0.0
        b.cavity(i).copy_(torch.tensor(f0 - w, device=dev).to(dtype))
    return b, states


def test_layout_offsets_are_aligned_and_disjoint(dev):
    b, _ = _batch(dev)
    end = 0
    for i, e in enumerate(EDGES):
        off = b.offset(i)
        assert off % 64 == 0 and off >= end
        end = off + e * 19 * ((e * e + 63) // 64 * 64)
    assert end == b.total_elements
    assert b.total_tiles == sum(e * ((e * e + 255) // 256) for e in EDGES)


def test_batch_matches_oracle_fp64(dev):
    b, states = _batch(dev)
    for _ in range(3):
        b.step()
        states = [ref.cavity_step(s, t, U_LID) for s, t in zip(states, TAUS)]
    for i in range(len(EDGES)):
        np.testing.assert_allclose(b.populations(i).cpu().numpy(), states[i], rtol=1e-12, atol=1e-14,
                                   err_msg=f"cavity {i}")


def _singles(dev, states, dtype):
    out = []
    for i, e in enumerate(EDGES):
        s = tl.FusedSolver3D(e, e, e, TAUS[i], dev, dtype, solid=wall_mask(e, dev), boundary="lid_y",
                             u_bc=(U_LID, 0.0, 0.0), layout="contiguous")
        out.append((s, torch.tensor(states[i], device=dev, dtype=dtype)))
    return out


def test_single_grid_lid_and_walls_match_oracle_fp64(dev):
    """The fused step with the wall mask and TLBM_BC_LID_Y is the oracle's cavity chain."""
    _, states = _batch(dev)
    singles = _singles(dev, states, torch.float64)
    for _ in range(3):
        singles = [(s, s.step(f)) for s, f in singles]
        states = [ref.cavity_step(x, t, U_LID) for x, t in zip(states, TAUS)]
    for i, (_, f) in enumerate(singles):
        np.testing.assert_allclose(f.cpu().numpy(), states[i], rtol=1e-12, atol=1e-14, err_msg=f"cavity {i}")


def test_batch_matches_single_grid_fused_step_fp32(dev):
    """Same arithmetic, different kernels: agreement to fp32 rounding (instruction selection may
    contract multiply-adds differently)."""
    b, states = _batch(dev, torch.float32)
    singles = _singles(dev, states, torch.float32)
    for _ in range(4):
        b.step()
        singles = [(s, s.step(f)) for s, f in singles]
    for i, (_, f) in enumerate(singles):
        d = float((b.populations(i) - f).abs().max())
        assert d < 2e-7, (i, d)


@pytest.mark.parametrize("grid", [1, 7, 64, 4096])
def test_persistent_grids_match_one_workgroup_per_tile(dev, grid):
    b0, _ = _batch(dev, torch.float32, seed=3)
    b1, _ = _batch(dev, torch.float32, seed=3)
    b1.grid = grid
    for _ in range(3):
        b0.step()
        b1.step()
    for i in range(len(EDGES)):
        assert torch.equal(b0.cavity(i), b1.cavity(i)), i


def test_graph_replays_match_eager_steps(dev):
    """Each replay of the captured ping-pong graph equals the same number of eager launches,
    with eager work and synchronisation between the replays."""
    be, _ = _batch(dev, torch.float32, seed=5)
    bg, _ = _batch(dev, torch.float32, seed=5)
    for rep in range(3):
        bg.run(6, graph=True, graph_steps=6)
        be.run(6, graph=False)
        torch.cuda.synchronize()
        assert all(torch.equal(bg.cavity(i), be.cavity(i)) for i in range(len(EDGES))), f"replay {rep}"
        _ = (bg.a * 2.0).sum().item()  # unrelated eager work between replays
    bg.run(9, graph=True, graph_steps=4)  # two graph launches + one eager step
    be.run(9, graph=False)
    torch.cuda.synchronize()
    assert all(torch.isfinite(be.cavity(i)).all() for i in range(len(EDGES)))
    assert all(torch.equal(bg.cavity(i), be.cavity(i)) for i in range(len(EDGES)))


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_deviation_storage_tracks_fp32(dev, dtype):
    b32, _ = _batch(dev, torch.float32, seed=7)
    b16, _ = _batch(dev, dtype, deviation=True, seed=7)
    for _ in range(20):
        b32.step()
        b16.step()
    for i in range(len(EDGES)):
        a, c = b16.populations(i).double(), b32.populations(i).double()
        assert torch.isfinite(a).all()
        # 16-bit rounding of f - w (|f - w| < 0.02 here) per step, 20 steps
        tol = 5e-4 if dtype == torch.float16 else 4e-3
        assert float((a - c).abs().max()) < tol, (i, float((a - c).abs().max()))


def test_lid_drives_the_flow(dev):
    edges, taus = [24, 17], [0.56, 0.6]
    b = BatchedCavities(edges, taus, u_lid=U_LID, device=dev, dtype=torch.float32)
    b.fill_equilibrium()
    b.run(300, graph=True, graph_steps=20)
    for i, e in enumerate(edges):
        rho, ux, uy, uz = tl.macroscopic3d(b.populations(i))
        inner = (slice(1, e - 1), slice(1, e - 1), slice(1, e - 1))
        top = ux[1:e - 1, e - 2, 1:e - 1].mean()
        assert float(top) > 0.2 * U_LID  # the layer under the lid is dragged along +x
        assert abs(float(rho[inner].mean()) - 1.0) < 1e-2
        assert float(uy[inner].abs().max()) > 1e-3  # a recirculating vortex, not a plain shear


def test_config_generator_matches_its_formula():
    edges, taus = cavity_batch_config(64, seed=0)
    assert len(edges) == 64 and min(edges) >= 32 and max(edges) <= 160
    re = 3 * 0.1 * np.array(edges) / (np.array(taus) - 0.5)
    assert re.min() >= 100 - 1e-9 and re.max() <= 1000 + 1e-9


def test_argument_errors(dev):
    with pytest.raises(ValueError):
        BatchedCavities([8, 2], [0.6, 0.6], device=dev)
    with pytest.raises(ValueError):
        BatchedCavities([8], [0.5], device=dev).step()
    with pytest.raises(ValueError):
        BatchedCavities([8] * 129, [0.6] * 129, device=dev)


# -- two steps per launch (tlbm_cavities_step2) -----------------------------------------------

EDGES2 = [8, 11, 13, 16, 5, 70]  # one x tile with clamped lanes .. a ragged second x tile
TAUS2 = [0.8, 0.55, 0.62, 0.56, 1.3, 0.57]


def _batch2(dev, dtype=torch.float32, deviation=False, seed=30, two_step=True):
    b = BatchedCavities(EDGES2, TAUS2, u_lid=U_LID, device=dev, dtype=dtype, deviation=deviation,
                        temporal_blocking=two_step)
    states = []
    for i, e in enumerate(EDGES2):
        f0 = F.random_state(19, (e, e, e), seed=seed + i, noise=5e-3, u_amp=0.03)
        states.append(f0)
        w = ref.D3Q19.w.reshape(19, 1, 1, 1) if deviation else 0.0
        b.cavity(i).copy_(torch.tensor(f0 - w, device=dev).to(dtype))
    return b, states


def test_two_step_batch_matches_oracle(dev):
    """Two launches of two steps = four steps of the oracle's cavity chain (fp32 storage vs fp64)."""
    b, states = _batch2(dev)
    assert b.two_step
    b.step2()
    b.step2()
    states = [ref.cavity_step(ref.cavity_step(ref.cavity_step(ref.cavity_step(s, t, U_LID), t, U_LID), t, U_LID),
                              t, U_LID) for s, t in zip(states, TAUS2)]
    for i in range(len(EDGES2)):
        np.testing.assert_allclose(b.populations(i).cpu().numpy(), states[i], rtol=1e-5, atol=2e-7,
                                   err_msg=f"cavity {i}")


def test_two_step_matches_one_step_launches(dev):
    """Same arithmetic in the one- and the two-step kernels, which are different code: instruction
    selection may round a value differently by an ulp (6e-8 at f ~ 0.5) in a step, and such
    differences accumulate over the 6 steps compared."""
    b2, _ = _batch2(dev, seed=31)
    b1, _ = _batch2(dev, seed=31, two_step=False)
    for _ in range(3):
        b2.step2()
        b1.step2()  # two one-step launches
    for i in range(len(EDGES2)):
        d = float((b2.populations(i) - b1.populations(i)).abs().max())
        assert d < 6 * 6e-8, (i, d)


def test_two_step_z_chunks_agree_bit_for_bit(dev):
    """The z chunk only changes which workgroup recomputes which halo plane, not any value."""
    out = []
    for zc in (0, 1, 3, 7):
        b, _ = _batch2(dev, seed=32)
        b.zc = zc
        b.step2()
        b.step2()
        out.append([b.cavity(i).clone() for i in range(len(EDGES2))])
    for other in out[1:]:
        assert all(torch.equal(x, y) for x, y in zip(out[0], other))


def test_two_step_graph_replays_match_eager_pairs(dev):
    be, _ = _batch2(dev, seed=33)
    bg, _ = _batch2(dev, seed=33)
    for rep in range(3):
        bg.run(8, graph=True, graph_steps=8)
        be.run(8, graph=False)
        torch.cuda.synchronize()
        assert all(torch.equal(bg.cavity(i), be.cavity(i)) for i in range(len(EDGES2))), f"replay {rep}"
        _ = (bg.a * 2.0).sum().item()  # unrelated eager work between replays
    bg.run(11, graph=True, graph_steps=4)  # two graph launches, one eager pair, one single step
    be.run(11, graph=False)
    torch.cuda.synchronize()
    assert all(torch.equal(bg.cavity(i), be.cavity(i)) for i in range(len(EDGES2)))
    with pytest.raises(ValueError):
        bg.run(12, graph=True, graph_steps=6)


def test_two_step_deviation_storage_tracks_fp32(dev):
    b32, _ = _batch2(dev, seed=34)
    b16, _ = _batch2(dev, torch.float16, deviation=True, seed=34)
    for _ in range(10):
        b32.step2()
        b16.step2()
    for i in range(len(EDGES2)):
        a, c = b16.populations(i).double(), b32.populations(i).double()
        assert torch.isfinite(a).all()
        assert float((a - c).abs().max()) < 5e-4, (i, float((a - c).abs().max()))


def test_autotune_keeps_the_fastest_shape(dev):
    b, _ = _batch2(dev, seed=35, two_step=False)
    before = [b.populations(i).clone() for i in range(len(EDGES2))]
    res = b.autotune(candidates=[0, 256], n_steps=4, zc_candidates=[0, 4])
    torch.cuda.synchronize()
    # tuning times steps on a copy: the state it leaves is the state it found
    assert all(torch.equal(b.populations(i), before[i]) for i in range(len(EDGES2)))
    assert set(res) == {"grid=0", "grid=256", "zc=0", "zc=4"}
    best = min(res, key=res.get)
    assert b.two_step == best.startswith("zc=")
    assert b.grid in (0, 256)
    b.run(6, graph=True, graph_steps=4)
    torch.cuda.synchronize()
    assert all(torch.isfinite(b.cavity(i)).all() for i in range(len(EDGES2)))


def test_two_step_argument_errors(dev):
    with pytest.raises(ValueError):
        BatchedCavities([8], [0.6], device=dev, dtype=torch.float64, temporal_blocking=True)
    from tensorlbm_amd import _native as nat

    b = BatchedCavities([8], [0.6], device=dev)
    rc = nat.lib().tlbm_cavities_step2(nat.store_code(torch.float64), b.a.data_ptr(), b.b.data_ptr(), b.a.numel(),
                                       b._table, 1, 0.1, 0, None)
    assert rc != 0
    rc = nat.lib().tlbm_cavities_step2(b._store, b.a.data_ptr(), b.b.data_ptr(), b.a.numel(), b._table, 1, 0.1, -1,
                                       None)
    assert rc == nat.ERR_INVALID
