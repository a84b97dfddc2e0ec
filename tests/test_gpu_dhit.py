"""Decaying homogeneous isotropic turbulence: reduced-precision storage against FP64, per field."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from tensorlbm_amd import diagnostics as dg
from tensorlbm_amd import fields
from tensorlbm_amd.closure import EddyViscosityMLP3D

pytestmark = pytest.mark.gpu

N, STEPS = 48, 150


def _init(lat, dtype, deviation, dev):
    u = fields.random_solenoidal_field(N, 4.0, 0.03, seed=7, device=dev, dtype=torch.float64)
    rho = torch.ones((N, N, N), device=dev, dtype=torch.float64)
    eq = tl.equilibrium27 if lat == "D3Q27" else tl.equilibrium3d
    f = eq(rho, u[0].contiguous(), u[1].contiguous(), u[2].contiguous())
    if deviation:
        q = f.shape[0]
        f = f - tl.lattice.BY_Q[q].W_EXACT64.to(dev).reshape(q, 1, 1, 1)
    return f.to(dtype)


def _run(lat, coll, tau, dtype, deviation, dev, closure=None):
    f = _init(lat, dtype, deviation, dev)
    s = tl.FusedSolver3D(N, N, N, tau, dev, dtype, lattice=lat, collision=coll, deviation=deviation,
                         closure=closure, C_s=0.17 if closure is not None else 0.0)
    for _ in range(STEPS):
        f = s.step(f)
    return f


@pytest.mark.parametrize("lat,coll,tau", [("D3Q27", "GEIER", 0.505), ("D3Q19", "BGK", 0.51)])
def test_deviation_storage_errors_vs_fp64(dev, lat, coll, tau):
    ref = _run(lat, coll, tau, torch.float64, False, dev)
    e32 = dg.field_errors(_run(lat, coll, tau, torch.float32, False, dev), ref)
    e16 = dg.field_errors(_run(lat, coll, tau, torch.float16, True, dev), ref, deviation=True)
    eb16 = dg.field_errors(_run(lat, coll, tau, torch.bfloat16, True, dev), ref, deviation=True)
    print(f"{lat} {coll} tau={tau}: fp32 {e32}\n fp16-dev {e16}\n bf16-dev {eb16}")
    assert e32["u"] < 5e-5 and e32["pi"] < 5e-4
    assert e16["u"] < 1e-3 and e16["pi"] < 5e-3
    assert eb16["u"] < 1e-2 and eb16["pi"] < 5e-2
    assert e32["u"] < e16["u"] < eb16["u"]


@pytest.mark.parametrize("operands", ["fp16", "bf16"])
def test_fp16_storage_with_learned_closure_vs_fp64(dev, operands):
    """The FP16-store + closure form stays within 1e-3 (velocity, relative L2) of its FP64 run,
    with either matrix-core operand form of the closure."""
    m = EddyViscosityMLP3D(dev, seed=3, operands=operands)
    m.set_feature_stats(torch.zeros(7), torch.full((7,), 1e-3))
    ref = _run("D3Q19", "BGK", 0.51, torch.float64, False, dev, closure=m)
    e32 = dg.field_errors(_run("D3Q19", "BGK", 0.51, torch.float32, False, dev, closure=m), ref)
    e16 = dg.field_errors(_run("D3Q19", "BGK", 0.51, torch.float16, True, dev, closure=m), ref, deviation=True)
    print(f"closure ({operands}): fp32 {e32}\n fp16-dev {e16}")
    # the closure's inputs carry ~22 bits and its layer-1 activations 11 (fp16): the rare
    # activation rounding flips between storage precisions stay far below the fp16-store error
    assert e32["u"] < 3e-4
    assert e16["u"] < 1e-3


def test_dhit_256_geier_per_field_errors_vs_fp64(dev):
    """The configuration at full size: 256³ decaying turbulence (k0 = 4, u_rms = 0.03), D3Q27 Geier cumulants,
    τ = 0.505, 200 steps; fp32 and 16-bit deviation storage against an FP64 run, per field."""
    n, steps = 256, 200
    u = fields.random_solenoidal_field(n, 4.0, 0.03, seed=11, device=dev, dtype=torch.float64)
    f0 = tl.equilibrium27(torch.ones((n, n, n), device=dev, dtype=torch.float64), u[0].contiguous(),
                          u[1].contiguous(), u[2].contiguous())
    del u
    w = tl.lattice.D3Q27.W_EXACT64.to(dev).reshape(27, 1, 1, 1)

    def run(dtype, deviation):
        s = tl.FusedSolver3D(n, n, n, 0.505, dev, dtype, lattice="D3Q27", collision="GEIER", deviation=deviation)
        f, _ = s.run((f0 - w if deviation else f0).to(dtype), steps)
        out = f.contiguous()
        del s, f
        return out

    ref = run(torch.float64, False)
    errs = {}
    for name, dtype, dv in (("fp32", torch.float32, False), ("fp16dev", torch.float16, True),
                            ("bf16dev", torch.bfloat16, True)):
        got = run(dtype, dv)
        errs[name] = dg.field_errors(got, ref, deviation=dv)
        del got
        torch.cuda.empty_cache()
    print({k: {f: f"{v:.2e}" for f, v in e.items()} for k, e in errs.items()})
    assert errs["fp32"]["u"] < 5e-5 and errs["fp32"]["pi"] < 5e-4
    assert errs["fp16dev"]["u"] < 1e-3 and errs["fp16dev"]["pi"] < 5e-3
    assert errs["bf16dev"]["u"] < 1e-2 and errs["bf16dev"]["pi"] < 5e-2
