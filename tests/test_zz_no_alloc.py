"""``tlbm_mlp_train_step_noalloc`` (partials in the library's own device array) against the
allocating pair ``tlbm_mlp_grad`` + ``tlbm_mlp_adam`` (partials in a caller buffer).

Both run the same kernels over the same block partition, so parameters, Adam moments, loss and
gradient must agree bit for bit, on two calls in a row, the second with a smaller minibatch.
"""

from __future__ import annotations

import ctypes

import pytest
import torch

from tensorlbm_amd import _native as nat
from tensorlbm_amd.closure import EddyViscosityMLP3D, sizes

pytestmark = pytest.mark.gpu

LR, B1, B2, EPS, CLIP = 1e-3, 0.9, 0.999, 1e-8, 1.0


def _data(n, seed, dev):
    g = torch.Generator().manual_seed(seed)
    x = (1e-3 * torch.randn(n, 7, generator=g)).to(dev)
    t = (1e-3 * torch.rand(n, generator=g)).to(dev)
    return x.contiguous(), t.contiguous()


def _allocating_step(lib, params, m, v, step, x, t, s):
    bb = sizes()["batch_block"]
    n = int(x.shape[0])
    nblk = (n + bb - 1) // bb
    partials = torch.empty((nblk, sizes()["nparam"] + 1), dtype=torch.float32, device=x.device)
    loss = torch.empty(1, dtype=torch.float32, device=x.device)
    grad = torch.empty(sizes()["nparam"], dtype=torch.float32, device=x.device)
    nat.check(lib.tlbm_mlp_grad(params.data_ptr(), x.data_ptr(), t.data_ptr(), n, partials.data_ptr(), s),
              "tlbm_mlp_grad")
    nat.check(lib.tlbm_mlp_adam(params.data_ptr(), partials.data_ptr(), nblk, m.data_ptr(), v.data_ptr(), step,
                                LR, B1, B2, EPS, CLIP, loss.data_ptr(), grad.data_ptr(), s), "tlbm_mlp_adam")
    return loss, grad


def _noalloc_step(lib, params, m, v, step, x, t, s, loss, grad):
    nat.check(lib.tlbm_mlp_train_step_noalloc(params.data_ptr(), x.data_ptr(), t.data_ptr(), int(x.shape[0]),
                                              m.data_ptr(), v.data_ptr(), step, LR, B1, B2, EPS, CLIP,
                                              loss.data_ptr(), grad.data_ptr(), s),
              "tlbm_mlp_train_step_noalloc")


def test_noalloc_training_step_matches_the_allocating_pair(dev):
    lib = nat.lib()
    model = EddyViscosityMLP3D(dev, seed=11)
    nparam = model.nparam
    pa = model.params.clone()
    pb = model.params.clone()
    ma, va = torch.zeros(nparam, device=dev), torch.zeros(nparam, device=dev)
    mb, vb = torch.zeros(nparam, device=dev), torch.zeros(nparam, device=dev)
    loss_b = torch.full((1,), float("nan"), device=dev)
    grad_b = torch.full((nparam,), float("nan"), device=dev)
    s = nat.stream_of(pa)
    # second call smaller than the first: the array keeps the first call's extra blocks, which
    # the second call must not read
    for step, (n, seed) in enumerate(((5000, 1), (1234, 2)), start=1):
        print(f"noalloc step {step}: n = {n}", flush=True)
        x, t = _data(n, seed, dev)
        loss_a, grad_a = _allocating_step(lib, pa, ma, va, step, x, t, s)
        _noalloc_step(lib, pb, mb, vb, step, x, t, s, loss_b, grad_b)
        torch.cuda.synchronize()
        assert torch.equal(loss_a, loss_b), (loss_a.item(), loss_b.item())
        assert torch.equal(grad_a, grad_b)
        assert torch.equal(pa, pb)
        assert torch.equal(ma, mb) and torch.equal(va, vb)
        assert torch.isfinite(pa).all()
    # the update moved the parameters
    assert not torch.equal(pa[:nparam], model.params[:nparam])


def test_noalloc_rejects_batches_beyond_its_array(dev):
    lib = nat.lib()
    limit = nat.MLP_NOALLOC_MAX_SAMPLES
    p = ctypes.c_void_p(0x1000)
    rc = lib.tlbm_mlp_train_step_noalloc(p, p, p, limit + 1, p, p, 1, LR, B1, B2, EPS, CLIP, None, None, None)
    assert rc == nat.ERR_SHAPE
    assert b"no-allocation limit" in lib.tlbm_last_error()
    # the largest batch it takes runs and agrees with the allocating pair
    model = EddyViscosityMLP3D(dev, seed=3)
    x, t = _data(limit, 9, dev)
    pa, pb = model.params.clone(), model.params.clone()
    z = [torch.zeros(model.nparam, device=dev) for _ in range(4)]
    loss_b = torch.empty(1, device=dev)
    grad_b = torch.empty(model.nparam, device=dev)
    s = nat.stream_of(pa)
    loss_a, grad_a = _allocating_step(lib, pa, z[0], z[1], 1, x, t, s)
    _noalloc_step(lib, pb, z[2], z[3], 1, x, t, s, loss_b, grad_b)
    torch.cuda.synchronize()
    assert torch.equal(pa, pb) and torch.equal(loss_a, loss_b) and torch.equal(grad_a, grad_b)
