"""Two steps per launch (``tlbm_step2``) against two single-step launches and the oracle.

The two-step kernel keeps the intermediate lattice in LDS, in the storage
format, and runs the same per-cell code as ``tlbm_step``, so its result equals
two single steps up to rounding: the two kernels inline that code into
different contexts and the compiler may fuse multiply-adds differently, which
moves a value by an ulp of fp32 (or, rounded into 16-bit storage, by one ulp of
the stored type).  The tests require that bar everywhere and identical values
almost everywhere (measured bit-identical here).  Grids are ragged against the
64 x 8 tiles (32 x 8 for D3Q27 fp32) and the z
blocking, and one is smaller than a tile so the halo wraps more than once.  The
integer canary pins the LDS addressing (including the z slots and bounce-back)
to the oracle's pull.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref
from tests import _fields as F

pytestmark = pytest.mark.gpu

LAT = {19: "D3Q19", 27: "D3Q27"}
STORES = {"fp32": (torch.float32, False), "fp16dev": (torch.float16, True), "bf16dev": (torch.bfloat16, True)}


def _solvers(q, shape, tau, store="fp32", **kw):
    dtype, dev_store = STORES[store]
    nz, ny, nx = shape
    one = tl.FusedSolver3D(nz, ny, nx, tau, "cuda:0", dtype, lattice=LAT[q], deviation=dev_store,
                           temporal_blocking=False, **kw)
    two = tl.FusedSolver3D(nz, ny, nx, tau, "cuda:0", dtype, lattice=LAT[q], deviation=dev_store,
                           temporal_blocking=True, **kw)
    return one, two


def _initial(q, shape, store, seed=3, u_amp=0.04):
    dtype, dev_store = STORES[store]
    f = F.random_state(q, shape, seed=seed, noise=2e-3, u_amp=u_amp)
    if dev_store:
        f = f - F.LATS[q].w.reshape(-1, 1, 1, 1)
    return torch.from_numpy(f).to("cuda:0", dtype=dtype)


def _pair(one, two, f0, pairs=2, force=False):
    fa = torch.zeros(3, dtype=torch.float64, device="cuda:0") if force else None
    fb = torch.zeros(3, dtype=torch.float64, device="cuda:0") if force else None
    a = b = f0
    for _ in range(pairs):
        a = one.step(one.step(a, force=fa), force=fa)
        b = two.step2(b, force=fb)
    torch.cuda.synchronize()
    return a, b, fa, fb


# relative size of one ulp of the stored type
_ULP = {torch.float32: 2.0**-23, torch.float16: 2.0**-10, torch.bfloat16: 2.0**-7}


def _assert_equal(a, b):
    """Equal up to a few ulps of the stored type, identical in at least 99.9 % of the values."""
    a32, b32 = a.double().cpu(), b.double().cpu()
    assert torch.isfinite(a32).all()
    diff = (a32 - b32).abs()
    bound = 4 * _ULP[a.dtype] * torch.maximum(a32.abs(), b32.abs()) + 1e-30
    assert bool((diff <= bound).all()), f"max|diff| {diff.max().item():.3e}"
    same = float((diff == 0).double().mean())
    assert same >= 0.999, f"only {same:.5f} of the values are identical"


@pytest.mark.parametrize("q,collision", [(19, "BGK"), (19, "MRT"), (19, "CM"), (19, "CUMULANT"),
                                         (27, "BGK"), (27, "MRT"), (27, "CM"), (27, "CUMULANT"), (27, "GEIER")])
@pytest.mark.parametrize("store", ["fp32", "fp16dev"])
def test_periodic_collisions_match_two_steps(q, collision, store):
    shape = (20, 14, 70)
    one, two = _solvers(q, shape, 0.62, store, collision=collision)
    assert two.two_step
    a, b, _, _ = _pair(one, two, _initial(q, shape, store))
    _assert_equal(a, b)


@pytest.mark.parametrize("store", ["fp32", "fp16dev", "bf16dev"])
@pytest.mark.parametrize("q", [19, 27])
def test_smagorinsky_and_storages(q, store):
    shape = (11, 9, 67)
    one, two = _solvers(q, shape, 0.53, store, C_s=0.17)
    assert two.two_step
    a, b, _, _ = _pair(one, two, _initial(q, shape, store, seed=5), pairs=3)
    _assert_equal(a, b)


@pytest.mark.parametrize("shape", [(5, 3, 4), (2, 7, 1), (1, 1, 9), (33, 1, 130)])
def test_grids_smaller_than_a_tile(shape):
    one, two = _solvers(19, shape, 0.7)
    a, b, _, _ = _pair(one, two, _initial(19, shape, "fp32", seed=9))
    _assert_equal(a, b)


@pytest.mark.parametrize("store", ["fp32", "fp16dev"])
def test_sphere_channel(store):
    nz, ny, nx = 18, 22, 70  # (nx - 1) % 64 = 5: outlet column pulls stay in the last tile
    solid = torch.from_numpy(ref.sphere_mask(nx, ny, nz, 20.0, 11.0, 9.0, 5.5)).to(torch.int8)
    one, two = _solvers(19, (nz, ny, nx), 0.56, store, solid=solid, boundary="channel_x", u_bc=(0.05, 0.0, 0.0),
                        C_s=0.17)
    assert two.two_step
    a, b, _, _ = _pair(one, two, _initial(19, (nz, ny, nx), store, seed=11, u_amp=0.02), pairs=3)
    _assert_equal(a, b)


def test_channel_walls_touching_inlet_and_outlet():
    """Solid cells on the inlet and outlet planes (channel walls): the step-one inlet / outlet
    lanes gather their neighbouring column, the wall cells among them bounce back."""
    nz, ny, nx = 10, 12, 70
    wall = np.zeros((nz, ny, nx), dtype=bool)
    wall[:, 0, :] = wall[:, -1, :] = True
    wall[3:6, 4:7, 30:34] = True
    one, two = _solvers(19, (nz, ny, nx), 0.6, "fp32", solid=torch.from_numpy(wall), boundary="channel_x",
                        u_bc=(0.04, 0.0, 0.0))
    assert two.two_step
    a, b, _, _ = _pair(one, two, _initial(19, (nz, ny, nx), "fp32", seed=18, u_amp=0.02), pairs=3)
    _assert_equal(a, b)


# Force parity bound.  The two kernels' populations agree to 4 ulps of the stored type and are
# identical in >= 99.9 % of the values, so the summed forces agree to the FP64 summation order
# plus those rare one-ulp flips; a cell counted twice or missed would move a component by its
# momentum rho u (~1e-2 here), far beyond these bounds.
_FORCE_ATOL = {"fp32": 1e-6, "fp16dev": 2e-3, "bf16dev": 2e-3}


@pytest.mark.parametrize("store", ["fp32", "fp16dev", "bf16dev"])
@pytest.mark.parametrize("collision", ["BGK", "MRT", "CM", "CUMULANT"])
def test_force_in_the_two_step_launch_channel(store, collision):
    """The Ladd force of both steps, sampled inside one two-step launch, equals the sum that two
    one-step launches add (channel + sphere; BGK takes the channel kernel, the others the
    run-time boundary kernel)."""
    nz, ny, nx = 12, 12, 70
    solid = torch.from_numpy(ref.sphere_mask(nx, ny, nz, 20.0, 6.0, 6.0, 3.5)).to(torch.int8)
    one, two = _solvers(19, (nz, ny, nx), 0.6, store, solid=solid, boundary="channel_x", u_bc=(0.05, 0.0, 0.0),
                        collision=collision)
    assert two.two_step
    f0 = _initial(19, (nz, ny, nx), store, seed=19, u_amp=0.02)
    a, b, fa, fb = _pair(one, two, f0, pairs=2, force=True)
    _assert_equal(a, b)
    fa, fb = fa.cpu().numpy(), fb.cpu().numpy()
    assert np.abs(fa).max() > 1e-2
    print(f"{store} {collision}: force {fa}, two-step - one-step {fb - fa}")
    np.testing.assert_allclose(fb, fa, rtol=0, atol=_FORCE_ATOL[store])


@pytest.mark.parametrize("shape", [(200, 20, 70), (9, 5, 30), (16, 9, 129)])
def test_force_counts_every_solid_cell_once(shape):
    """Scattered solid cells (8 %) on tile edges, tile halos, z-chunk boundaries and, on the
    30-wide grid, halo columns that wrap around a tile wider than the lattice: every solid cell
    is sampled once per step.  Periodic BGK takes the run-time boundary kernel."""
    nz, ny, nx = shape
    rng = np.random.default_rng(24)
    solid = torch.from_numpy(rng.random(shape) < 0.08).to(torch.int8)
    one, two = _solvers(19, shape, 0.7, "fp32", solid=solid)
    assert two.two_step
    a, b, fa, fb = _pair(one, two, _initial(19, shape, "fp32", seed=20, u_amp=0.03), pairs=2, force=True)
    _assert_equal(a, b)
    fa, fb = fa.cpu().numpy(), fb.cpu().numpy()
    assert np.abs(fa).max() > 1e-2
    np.testing.assert_allclose(fb, fa, rtol=0, atol=_FORCE_ATOL["fp32"])


def test_two_step_force_matches_two_oracle_steps():
    """One two-step launch against two steps of the oracle's obstacle step (fp64): final
    populations and the sum of the two steps' Ladd forces."""
    from tensorlbm_amd import _native as nat

    nz, ny, nx = 10, 12, 70
    solid = ref.sphere_mask(nx, ny, nz, 30.0, 6.0, 5.0, 3.5)
    f0 = F.random_state(19, (nz, ny, nx), seed=25, noise=2e-3, u_amp=0.03)
    want1, force1 = ref.fused_obstacle_step(f0, 0.62, solid, with_force=True)
    want, force2 = ref.fused_obstacle_step(want1, 0.62, solid, with_force=True)
    two = tl.FusedSolver3D(nz, ny, nx, 0.62, "cuda:0", torch.float32, solid=torch.from_numpy(solid))
    assert two.two_step
    acc = torch.zeros(3, dtype=torch.float64, device="cuda:0")
    got = two.step2(torch.from_numpy(f0).to("cuda:0", torch.float32), force=acc).cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(acc.cpu().numpy(), np.add(force1, force2), rtol=0, atol=1e-5)
    # D3Q27 declines a force buffer in the two-step launch (its solver takes two one-step launches)
    coll = nat.make_collision(nat.COLL_BGK, 0.6)
    f27 = torch.zeros((27, 4, 4, 64), device="cuda:0")
    out = torch.empty_like(f27)
    m = torch.ones((4, 4, 64), dtype=torch.int8, device="cuda:0")
    rc = nat.lib().tlbm_step2(27, nat.STORE_F32, f27.data_ptr(), None, out.data_ptr(), None, 4, 4, 64, coll,
                              m.data_ptr(), None, acc.data_ptr(), None)
    assert rc == nat.ERR_UNSUPPORTED
    assert b"D3Q19" in nat.lib().tlbm_last_error()
    rc = nat.lib().tlbm_step2(19, nat.STORE_F32, f27.data_ptr(), None, out.data_ptr(), None, 4, 4, 64, coll,
                              None, None, acc.data_ptr(), None)
    assert rc == nat.ERR_INVALID


def test_d3q27_tile_channel():
    nz, ny, nx = 10, 12, 40  # 32 x 8 tiles for D3Q27 fp32: (nx - 1) % 32 = 7
    solid = torch.from_numpy(ref.sphere_mask(nx, ny, nz, 14.0, 6.0, 5.0, 3.0)).to(torch.int8)
    one, two = _solvers(27, (nz, ny, nx), 0.6, "fp32", solid=solid, boundary="channel_x", u_bc=(0.04, 0.0, 0.0),
                        collision="GEIER")
    assert two.two_step
    a, b, _, _ = _pair(one, two, _initial(27, (nz, ny, nx), "fp32", seed=12, u_amp=0.02))
    _assert_equal(a, b)


@pytest.mark.parametrize("config", [None, {"far_field_faces": ["y-", "y+"], "periodic_faces": ["z-", "z+"]},
                                    {"far_field_faces": ["z-", "y+"], "periodic_faces": ["x-", "x+"]}])
def test_far_field(config):
    nz, ny, nx = 12, 13, 69
    solid = torch.from_numpy(ref.sphere_mask(nx, ny, nz, 22.0, 6.0, 6.0, 3.5)).to(torch.int8)
    one, two = _solvers(19, (nz, ny, nx), 0.58, "fp32", solid=solid, boundary="far_field", u_bc=(0.05, 0.01, 0.0),
                        bc_config=config)
    a, b, _, _ = _pair(one, two, _initial(19, (nz, ny, nx), "fp32", seed=13, u_amp=0.02))
    _assert_equal(a, b)


def test_lid_cavity():
    n = 20
    wall = np.zeros((n, n, n), dtype=bool)
    wall[:, :, 0] = wall[:, :, -1] = wall[:, 0, :] = wall[0] = wall[-1] = True
    one, two = _solvers(19, (n, n, n), 0.6, "fp32", solid=torch.from_numpy(wall), boundary="lid_y",
                        u_bc=(0.08, 0.0, 0.0))
    a, b, _, _ = _pair(one, two, _initial(19, (n, n, n), "fp32", seed=14, u_amp=0.0), pairs=4)
    _assert_equal(a, b)


def test_outlet_straddling_tiles_falls_back():
    nz, ny, nx = 6, 6, 65  # (nx - 1) % 64 = 0: the outlet column's neighbours sit in another tile
    one, two = _solvers(19, (nz, ny, nx), 0.6, "fp32", boundary="channel_x", u_bc=(0.05, 0.0, 0.0))
    assert not two.two_step
    a, b, _, _ = _pair(one, two, _initial(19, (nz, ny, nx), "fp32", seed=15, u_amp=0.02))
    _assert_equal(a, b)


def test_run_uses_pairs_and_odd_tail():
    shape = (16, 12, 64)
    one, two = _solvers(19, shape, 0.6)
    f0 = _initial(19, shape, "fp32", seed=16)
    a, _ = one.run(f0, 5)
    b, _ = two.run(f0, 5)
    torch.cuda.synchronize()
    _assert_equal(a, b)


def test_canary_addressing_matches_oracle_pull():
    nz, ny, nx = 12, 14, 62
    solid = ref.sphere_mask(nx, ny, nz, 30.0, 7.0, 6.0, 3.5)
    fc = F.canary(19, nz, ny, nx)
    two = tl.FusedSolver3D(nz, ny, nx, 1e9, "cuda:0", torch.float32, solid=torch.from_numpy(solid))
    assert two.two_step
    got = two.step2(torch.from_numpy(fc).to("cuda:0", torch.float32)).cpu().numpy()
    want = ref.fullway_pull(ref.fullway_pull(fc, solid), solid)
    assert int((got.round().astype(np.int64) != want.round().astype(np.int64)).sum()) == 0


def test_two_fused_bgk_steps_match_oracle():
    shape = (10, 9, 70)
    f0 = F.random_state(19, shape, seed=17).astype(np.float32)
    two = tl.FusedSolver3D(*shape, 0.7, "cuda:0", torch.float32)
    got = two.step2(torch.from_numpy(f0).to("cuda:0")).cpu().numpy()
    want = ref.fused_step(ref.fused_step(f0.astype(np.float64), 0.7), 0.7)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)


def _headline_state(n=256, seed=21):
    """256³ channel + sphere (radius 24 at x = 64, y = z = 128): equilibrium of u = (0.05, 0, 0) with a seeded
    1e-3 velocity perturbation, built on the GPU."""
    dev = "cuda:0"
    a = torch.arange(n, dtype=torch.float32, device=dev)
    z, y, x = torch.meshgrid(a, a, a, indexing="ij")
    solid = ((x - 64.0) ** 2 + (y - n / 2) ** 2 + (z - n / 2) ** 2) <= 24.0**2
    del x, y, z
    g = torch.Generator(device="cpu").manual_seed(seed)
    pert = ((torch.rand((3, n, n, n), generator=g) * 2 - 1) * 1e-3).to(dev)
    one = torch.ones((n, n, n), device=dev)
    f = tl.equilibrium3d(one, 0.05 + pert[0], pert[1].contiguous(), pert[2].contiguous())
    return solid.to(torch.int8), f


def test_headline_workload_two_steps_match_single_steps():
    """The benchmarked launch itself (256³, 256 workgroups of 128 planes): one two-step launch = two steps."""
    n = 256
    solid, f0 = _headline_state(n)
    kw = dict(solid=solid, boundary="channel_x", u_bc=(0.05, 0.0, 0.0))
    one = tl.FusedSolver3D(n, n, n, 0.572, "cuda:0", torch.float32, temporal_blocking=False, **kw)
    two = tl.FusedSolver3D(n, n, n, 0.572, "cuda:0", torch.float32, temporal_blocking=True, **kw)
    assert two.two_step
    a, b, _, _ = _pair(one, two, f0, pairs=2)
    _assert_equal(a, b)


def test_full_size_periodic_mass_conserved():
    """256³ periodic BGK, 10 two-step launches: the FP64 sum of the populations moves by fp32 rounding only."""
    n = 256
    _, f0 = _headline_state(n, seed=22)
    two = tl.FusedSolver3D(n, n, n, 0.6, "cuda:0", torch.float32)
    assert two.two_step
    m0 = tl.moment_sums(f0)[0].item()
    f = f0
    for _ in range(10):
        f = two.step2(f)
    m1 = tl.moment_sums(f)[0].item()
    assert abs(m1 - m0) <= 1e-6 * m0, (m0, m1)


def test_truncated_last_z_chunk():
    """200 planes in chunks of 7 (the launch-shape policy for 8 columns on 256 CUs): the last chunk has 4."""
    shape = (200, 64, 64)
    solid = torch.from_numpy(ref.sphere_mask(64, 64, 200, 20.0, 32.0, 100.0, 9.5)).to(torch.int8)
    one, two = _solvers(19, shape, 0.58, solid=solid, boundary="channel_x", u_bc=(0.04, 0.0, 0.0))
    assert two.two_step
    a, b, _, _ = _pair(one, two, _initial(19, shape, "fp32", seed=23, u_amp=0.02), pairs=2)
    _assert_equal(a, b)


def test_default_policy():
    """Two steps per launch by default only for D3Q19 with fp32 storage."""
    assert tl.FusedSolver3D(8, 8, 64, 0.6, "cuda:0", torch.float32).two_step
    assert tl.FusedSolver3D(8, 8, 64, 0.6, "cuda:0", torch.float32, collision="MRT").two_step
    assert not tl.FusedSolver3D(8, 8, 64, 0.6, "cuda:0", torch.float16, deviation=True).two_step
    assert not tl.FusedSolver3D(8, 8, 64, 0.6, "cuda:0", torch.float32, lattice="D3Q27").two_step
    assert not tl.FusedSolver3D(8, 8, 64, 0.6, "cuda:0", torch.float64).two_step


def test_unsupported_arguments_raise():
    from tensorlbm_amd import _native as nat

    f = torch.zeros((19, 4, 4, 4), device="cuda:0", dtype=torch.float64)
    out = torch.empty_like(f)
    coll = nat.make_collision(nat.COLL_BGK, 0.6)
    rc = nat.lib().tlbm_step2(19, nat.store_code(torch.float64), f.data_ptr(), None, out.data_ptr(), None, 4, 4, 4,
                              coll, None, None, None, None)
    assert rc == nat.ERR_UNSUPPORTED
    assert b"float32" in nat.lib().tlbm_last_error()
