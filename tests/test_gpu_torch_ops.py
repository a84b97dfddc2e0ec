"""torch.ops.tlbm on the GPU: the dispatcher ops equal the ctypes path bit for bit (same kernels),
trace under torch.compile, replay from a CUDA graph and show up in the profiler."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref
from tensorlbm_amd import _native as nat
from tensorlbm_amd import torch_ops
from tests import _fields as F

pytestmark = pytest.mark.gpu


def _state(q, shape, seed=1, dtype=torch.float32):
    return torch.from_numpy(F.random_state(q, shape, seed=seed)).to("cuda:0", dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_eager_ops_equal_the_ctypes_path(dtype):
    t = torch_ops.load()
    f = _state(19, (6, 7, 9), dtype=dtype)
    rho, ux, uy, uz = tl.macroscopic3d(f)
    got = t.macroscopic(f)
    for a, b in zip(got, (rho, ux, uy, uz)):
        assert torch.equal(a, b)
    assert torch.equal(t.equilibrium(19, rho, ux, uy, uz), tl.equilibrium3d(rho, ux, uy, uz))
    assert torch.equal(t.stream(f), tl.stream3d(f))
    assert torch.equal(t.collide(f, nat.COLL_BGK, 0.62), tl.collide_bgk3d(f, 0.62))
    assert torch.equal(t.collide(f, nat.COLL_MRT, 0.62, [1.19, 1.4, 1.2, 1.19]),
                       tl.collide_mrt3d(f, 0.62, 1.19, 1.4, 1.2))
    f9 = _state(9, (8, 10), dtype=dtype)
    assert torch.equal(t.stream(f9), tl.stream(f9))
    r9, x9, y9 = t.macroscopic(f9)
    assert torch.equal(t.equilibrium(9, r9, x9, y9), tl.equilibrium(r9, x9, y9))


def test_fused_step_with_force_equals_the_solver():
    t = torch_ops.load()
    nz, ny, nx = 10, 12, 70
    solid = torch.from_numpy(ref.sphere_mask(nx, ny, nz, 20.0, 6.0, 5.0, 3.5))
    f0 = _state(19, (nz, ny, nx), seed=3)
    kw = dict(solid=solid, boundary="channel_x", u_bc=(0.05, 0.0, 0.0))
    for steps in (1, 2):
        s = tl.FusedSolver3D(nz, ny, nx, 0.6, "cuda:0", torch.float32, layout="contiguous",
                             temporal_blocking=steps == 2, **kw)
        fa = torch.zeros(3, dtype=torch.float64, device="cuda:0")
        fb = torch.zeros(3, dtype=torch.float64, device="cuda:0")
        want = s.step(f0, force=fa) if steps == 1 else s.step2(f0, force=fa)
        got = t.step(f0, nat.COLL_BGK, 0.6, solid=solid.to("cuda:0"), bc_kind=nat.BC_CHANNEL_X,
                     bc_u=[0.05, 0.0, 0.0], force=fb, steps=steps)
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        np.testing.assert_allclose(fb.cpu().numpy(), fa.cpu().numpy(), rtol=1e-12, atol=1e-15)


def test_deviation_storage_step_equals_the_solver():
    t = torch_ops.load()
    shape = (8, 9, 64)
    f0 = (_state(19, shape, seed=4) - torch.from_numpy(F.LATS[19].w).to("cuda:0", torch.float32)
          .reshape(19, 1, 1, 1)).to(torch.float16)
    s = tl.FusedSolver3D(*shape, 0.58, "cuda:0", torch.float16, deviation=True, layout="contiguous", C_s=0.17)
    want = s.step(f0)
    got = t.step(f0, nat.COLL_BGK, 0.58, tau_mode=nat.TAU_SMAG, cs_delta_sq=0.17**2, deviation=True)
    assert torch.equal(got, want)


def test_errors_raise_the_reference_types():
    t = torch_ops.load()
    f = _state(19, (4, 4, 4))
    with pytest.raises(ValueError, match="Unsupported dtype"):
        t.stream(f.to(torch.int32))
    with pytest.raises(ValueError):
        t.step(f, nat.COLL_BGK, 0.4)  # tau <= 0.5
    with pytest.raises(NotImplementedError):
        t.step(_state(9, (8, 8)), nat.COLL_BGK, 0.6, steps=2)
    with pytest.raises(ValueError, match="force"):
        t.step(f, nat.COLL_BGK, 0.6, solid=torch.zeros((4, 4, 4), dtype=torch.bool, device="cuda:0"),
               force=torch.zeros(3, device="cuda:0"))


def test_torch_compile_traces_the_ops():
    t = torch_ops.load()
    f0 = _state(19, (6, 8, 64), seed=5)

    def two_steps(f):
        g = t.step(f, nat.COLL_BGK, 0.6)
        return t.step(g, nat.COLL_BGK, 0.6) * 1.0

    compiled = torch.compile(two_steps, backend="aot_eager", fullgraph=True)
    assert torch.equal(compiled(f0), two_steps(f0))


def test_cuda_graph_replays_match_eager():
    t = torch_ops.load()
    f0 = _state(19, (8, 8, 64), seed=6)
    static_in = f0.clone()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        for _ in range(2):  # warm-up outside capture
            t.step(static_in, nat.COLL_BGK, 0.6, steps=2)
    torch.cuda.current_stream().wait_stream(stream)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        static_out = t.step(static_in, nat.COLL_BGK, 0.6, steps=2)
    f = f0
    for _ in range(3):
        static_in.copy_(f)
        g.replay()
        torch.cuda.synchronize()
        want = t.step(f, nat.COLL_BGK, 0.6, steps=2)
        torch.cuda.synchronize()
        assert torch.equal(static_out, want)
        f = want


def test_profiler_sees_the_op():
    t = torch_ops.load()
    f0 = _state(19, (4, 6, 64), seed=7)
    with torch.profiler.profile(activities=[torch.profiler.ProfilerActivity.CPU]) as prof:
        t.step(f0, nat.COLL_BGK, 0.6)
        torch.cuda.synchronize()
    assert any(e.key == "tlbm::step" for e in prof.key_averages())
