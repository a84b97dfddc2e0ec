"""The reference's own unit tests of the per-step ops, on the HIP kernels.

Each block mirrors one reference test file on its own inputs and, where the reference's bound is loose
(1e-4 – 2e-5 for what the arithmetic fixes far tighter), also asserts the tight bound of the kernel:

* tests/test_d2q9.py, tests/test_d3q19.py, tests/test_d3q27.py (lattice tables, equilibrium ↔ moments
  round trips, shape checks, D3Q27 stream / BGK / MRT / Smagorinsky BGK and MRT / bounce-back / Ladd
  force / mass correction; the D3Q27 Zou–He channel and its config are not part of this package);
* tests/core/test_lattice_contracts.py (opposites are involutive negations, weights sum to 1);
* tests/test_low_memory_d3q19.py (the low-memory moments equal the vector form);
* tests/test_d3q19_d3q27_mrt_consistency.py (fixed point, local conservation, bitwise repeatability);
* tests/test_cumulant_d3q19.py (fixed points, conservation, τ ordering, the SGS options and their errors);
* tests/test_cumulant_d3q27_geier.py (fixed point for any rates, conservation, damping; the transform
  identities are oracle pins on the CPU);
* tests/test_d3q27_cumulant_reference_invariants.py (legacy D3Q27: conserved raw moments, equilibrium
  fixed, with and without the domain-mean C_s);
* tests/test_correct_mass.py (D2Q9 / D3Q19 / D3Q27 rescaling).

Inputs are built on the GPU with the library's own equilibrium, as the reference builds them with its
own; random fields come from a seeded CPU generator.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import tensorlbm_amd as tl
from oracle import ref

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _full(shape, v, dtype=torch.float32):
    return torch.full(shape, float(v), dtype=dtype, device=DEV)


def _rand(shape, seed, dtype=torch.float32):
    g = torch.Generator().manual_seed(seed)
    return torch.rand(shape, generator=g, dtype=dtype).to(DEV)


def _raw_moments(f):
    """m000, m100, m010, m001 of populations, formed from the velocity table, not from the library's moments."""
    c = tl.lattice.BY_Q[int(f.shape[0])].C.to(device=f.device, dtype=f.dtype)
    flat = f.reshape(f.shape[0], -1)
    return torch.cat((flat.sum(0, keepdim=True), c.T @ flat), 0)


# ---------------------------------------------------------------------------
# tables (test_d2q9.py, test_d3q19.py, test_d3q27.py, core/test_lattice_contracts.py)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("c,w,opp,q,d", [(tl.C, tl.W, tl.OPPOSITE, 9, 2), (tl.C3D, tl.W3D, tl.OPPOSITE3D, 19, 3),
                                         (tl.C27, tl.W27, tl.OPPOSITE27, 27, 3)])
def test_tables(c, w, opp, q, d):
    assert tuple(c.shape) == (q, d) and tuple(w.shape) == (q,) and w.dtype == torch.float32
    assert abs(float(w.sum()) - 1.0) < 1e-6
    assert torch.equal(opp[opp], torch.arange(q))
    assert torch.equal(c[opp], -c)


@pytest.mark.parametrize("tab,eq", [(tl.lattice.D3Q19, tl.equilibrium3d), (tl.lattice.D3Q27, tl.equilibrium27)])
def test_float64_equilibrium_uses_unrounded_weights(tab, eq):
    rho = torch.ones((3, 4, 3), dtype=torch.float64, device=DEV)
    z = torch.zeros_like(rho)
    f = eq(rho, z, z, z)
    assert torch.equal(f[:, 0, 0, 0].cpu(), tab.W_EXACT64)
    assert torch.equal(tab.W_EXACT64.float(), tab.W) and not torch.equal(tab.W.double(), tab.W_EXACT64)


# ---------------------------------------------------------------------------
# equilibrium <-> moments
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("u", [(0.0, 0.0, 0.0), (0.05, -0.02, 0.0), (0.05, 0.02, -0.01)])
@pytest.mark.parametrize("q", [9, 19, 27])
def test_equilibrium_round_trip(q, u):
    shape = (6, 8) if q == 9 else (4, 6, 8)
    rho = torch.ones(shape, device=DEV)
    vel = [_full(shape, v) for v in u[: 2 if q == 9 else 3]]
    eq, mac = {9: (tl.equilibrium, tl.macroscopic), 19: (tl.equilibrium3d, tl.macroscopic3d),
               27: (tl.equilibrium27, tl.macroscopic27)}[q]
    f = eq(rho, *vel)
    assert tuple(f.shape) == (q, *shape)
    got = mac(f)
    bound = 1e-6 if not any(u) else 1e-5  # the reference's bounds
    for g, want in zip(got, (rho, *vel)):
        assert torch.allclose(g, want, atol=bound)
        assert float((g - want).abs().max()) < 1e-6  # what fp32 arithmetic gives


@pytest.mark.parametrize("q", [9, 19, 27])
def test_equilibrium_shape_mismatch_raises(q):
    shape = (6, 8) if q == 9 else (4, 6, 8)
    rho = torch.ones(shape, device=DEV)
    bad = torch.zeros(shape[:-1] + (shape[-1] - 1,), device=DEV)
    eq = {9: tl.equilibrium, 19: tl.equilibrium3d, 27: tl.equilibrium27}[q]
    args = [rho, torch.zeros_like(rho), bad] + ([] if q == 9 else [torch.zeros_like(rho)])
    with pytest.raises(ValueError, match="shapes must match"):
        eq(*args)


@pytest.mark.parametrize("dtype,tol", [(torch.float32, 2e-7), (torch.float64, 3e-16)])
def test_low_memory_moments_match_vector_form(dtype, tol):
    state = 0.02 + 0.1 * _rand((19, 4, 5, 6), 20260802, dtype)
    for a, b in zip(tl.macroscopic3d_low_memory(state), tl.macroscopic3d(state)):
        torch.testing.assert_close(a, b, rtol=tol, atol=tol)
    with pytest.raises(ValueError, match="shape"):
        tl.macroscopic3d_low_memory(torch.zeros((27, 2, 2, 2), device=DEV))


# ---------------------------------------------------------------------------
# D3Q27 stream, BGK, MRT, Smagorinsky, bounce-back, force (test_d3q27.py)
# ---------------------------------------------------------------------------


def _feq27(shape=(4, 5, 6), seed=0, umax=0.04, dtype=torch.float32):
    rho = _rand(shape, seed, dtype) + 0.5
    u = [_rand(shape, seed + 1 + i, dtype) * umax for i in range(3)]
    return rho, u, tl.equilibrium27(rho, *u)


def test_stream27_shape_mass_and_repeatability():
    _, _, f = _feq27()
    a, b = tl.stream27(f), tl.stream27(f)
    assert a.shape == f.shape and torch.equal(a, b)
    assert abs(float(a.double().sum()) - float(f.double().sum())) < 1e-12 * float(f.double().sum())


@pytest.mark.parametrize("name", ["collide_bgk27", "collide_mrt27", "collide_smagorinsky_bgk27",
                                  "collide_smagorinsky_mrt27"])
def test_d3q27_collisions_conserve_and_fix_equilibrium(name):
    op = getattr(tl, name)
    rho, u, f = _feq27()
    out = op(f, tau=0.7)
    assert out.shape == f.shape and bool(torch.isfinite(out).all())
    got = tl.macroscopic27(out)
    for g, want in zip(got, (rho, *u)):
        assert float((g - want).abs().max()) < 2e-6  # reference bound 1e-4 / 1e-5
    r1 = torch.ones((4, 5, 6), device=DEV)
    feq = tl.equilibrium27(r1, _full(r1.shape, 0.03), _full(r1.shape, 0.01), _full(r1.shape, -0.01))
    assert float((op(feq, tau=0.7) - feq).abs().max()) < 1e-6  # reference bound 1e-4


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_smagorinsky_mrt27_matches_oracle(dtype):
    rng = np.random.default_rng(27)
    f = ref.equilibrium(ref.D3Q27, 1.0 + 0.05 * rng.random((5, 6, 7)), *(0.04 * rng.random((3, 5, 6, 7))))
    f = f + 5e-2 * rng.standard_normal(f.shape) * ref.D3Q27.w.reshape(-1, 1, 1, 1)
    got = tl.collide_smagorinsky_mrt27(torch.from_numpy(f).to(DEV, dtype), 0.58, C_s=0.2, s_pi=1.3)
    want = ref.collide_smagorinsky_mrt(f, 0.58, 0.2, s_pi=1.3)
    plain = ref.collide_mrt(f, 0.58, s_pi=1.3)
    assert np.abs(want - plain).max() > 1e-5  # the closure acts on this state
    tol = dict(rtol=1e-5, atol=2e-7) if dtype == torch.float32 else dict(rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(got.double().cpu().numpy(), want, **tol)


def test_bounce_back27_and_force27():
    _, _, f = _feq27((4, 6, 10))
    mask = torch.zeros((4, 6, 10), dtype=torch.bool, device=DEV)
    assert all(float(v) == 0.0 for v in tl.compute_obstacle_forces_27(f, mask))
    mask[0] = True
    out = tl.bounce_back_cells_27(f, mask)
    assert out.shape == f.shape
    assert torch.equal(out[:, 0], f[tl.OPPOSITE27.to(DEV)][:, 0]) and torch.equal(out[:, 1:], f[:, 1:])
    rho = torch.ones((4, 6, 10), device=DEV)
    fs = tl.stream27(tl.equilibrium27(rho, _full(rho.shape, 0.05), torch.zeros_like(rho), torch.zeros_like(rho)))
    m = torch.zeros_like(rho, dtype=torch.bool)
    m[2, 3, 5] = True
    got = [float(v) for v in tl.compute_obstacle_forces_27(fs, m)]
    assert all(math.isfinite(v) for v in got)
    want = ref.obstacle_force(fs.double().cpu().numpy(), m.cpu().numpy())
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-9)
    with pytest.raises(ValueError, match="27"):
        tl.compute_obstacle_forces_27(torch.zeros((19, 4, 6, 10), device=DEV), m)


# ---------------------------------------------------------------------------
# MRT consistency (test_d3q19_d3q27_mrt_consistency.py)
# ---------------------------------------------------------------------------


def _mrt_case(q):
    g = torch.Generator().manual_seed(2718)
    shape = (2, 3, 4)
    rho = 0.9 + 0.2 * torch.rand(shape, generator=g)
    u = [-0.035 + 0.07 * torch.rand(shape, generator=g) for _ in range(3)]
    eq = tl.equilibrium3d if q == 19 else tl.equilibrium27
    feq = eq(rho.to(DEV), *(v.to(DEV) for v in u))
    pert = torch.zeros_like(feq)
    pert[0], pert[1:5] = 2.0e-4, -5.0e-5  # zero density and momentum for both stencils' first five slots
    return feq, feq + pert


@pytest.mark.parametrize("q", [19, 27])
def test_mrt_fixed_point_conservation_and_repeatability(q):
    op = tl.collide_mrt3d if q == 19 else tl.collide_mrt27
    feq, f = _mrt_case(q)
    out = op(feq, tau=0.8)
    assert out.dtype == torch.float32 and tuple(out.shape) == (q, 2, 3, 4)
    torch.testing.assert_close(out, feq, rtol=0.0, atol=1.0e-6)
    a = op(f.clone(), tau=0.8)
    torch.testing.assert_close(_raw_moments(a), _raw_moments(f), rtol=0.0, atol=1.0e-6)
    assert torch.equal(a, op(f.clone(), tau=0.8))
    # the reference's D3Q27 MRT caches float32 matrices and refuses float64; here float64 is computed
    # in float64 and agrees with the float32 result to float32 rounding
    a64 = op(f.double(), tau=0.8)
    assert a64.dtype == torch.float64 and float((a64 - a.double()).abs().max()) < 1e-7


# ---------------------------------------------------------------------------
# D3Q19 cumulant (test_cumulant_d3q19.py)
# ---------------------------------------------------------------------------


def _moving_eq19(shape=(2, 3, 4)):
    rho = torch.ones(shape, device=DEV)
    return tl.equilibrium3d(rho, _full(shape, 0.03), _full(shape, -0.02), _full(shape, 0.01))


def test_cumulant19_fixed_points_and_conservation():
    rho = torch.ones((2, 3, 4), device=DEV)
    z = torch.zeros_like(rho)
    for feq in (tl.equilibrium3d(rho, z, z, z), _moving_eq19()):
        assert float((tl.collide_cumulant_d3q19(feq, tau=0.8) - feq).abs().max()) < 1e-7  # reference 2e-5
    f = _moving_eq19() + 1e-4 * torch.linspace(-1, 1, 19, device=DEV).view(-1, 1, 1, 1)
    out = tl.collide_cumulant_d3q19(f, tau=0.8)
    for a, b in zip(tl.macroscopic3d(out), tl.macroscopic3d(f)):
        assert float((a - b).abs().max()) < 4e-7  # a few ulp of ρ ≈ 1; reference 2e-5
    assert out.shape == f.shape and bool(torch.isfinite(out).all())


def test_cumulant19_higher_tau_relaxes_less_and_smagorinsky_is_finite():
    feq = _moving_eq19()
    f = feq + 1e-4 * torch.linspace(-1, 1, 19, device=DEV).view(-1, 1, 1, 1)
    low = (tl.collide_cumulant_d3q19(f, tau=0.55) - feq).abs().max()
    high = (tl.collide_cumulant_d3q19(f, tau=5.0) - feq).abs().max()
    assert high > low
    out = tl.collide_cumulant_d3q19(f, tau=0.55, C_s=0.1)
    assert out.shape == f.shape and bool(torch.isfinite(out).all())


@pytest.mark.parametrize("name,coef", [("C_w", 0.5), ("C_v", 0.025)])
def test_cumulant19_gradient_sgs_conserves(name, coef):
    z, y, x = torch.meshgrid(*(torch.arange(n, dtype=torch.float64, device=DEV) for n in (4, 5, 6)), indexing="ij")
    f = tl.equilibrium3d(torch.ones_like(x), 0.01 + 0.001 * y, -0.005 + 0.0005 * z, 0.0003 * x)
    f[1] += 1.0e-4
    f[2] -= 1.0e-4
    out = tl.collide_cumulant_d3q19(f, tau=0.55, **{name: coef})
    assert bool(torch.isfinite(out).all())
    for a, b in zip(tl.macroscopic3d(out), tl.macroscopic3d(f)):
        torch.testing.assert_close(a, b, atol=2.0e-8, rtol=0.0)
    te = tl.gradient_closure_tau(f, 0.55, **{name: coef})
    assert bool((te >= 0.55).all()) and bool((te > 0.55).any())


def test_cumulant19_sgs_argument_errors():
    rho = torch.ones((4, 5, 6), device=DEV)
    z = torch.zeros_like(rho)
    f = tl.equilibrium3d(rho, z, z, z)
    with pytest.raises(ValueError, match="only one"):
        tl.collide_cumulant_d3q19(f, tau=0.55, C_s=0.1, C_w=0.5)
    with pytest.raises(ValueError, match="solid_mask"):
        tl.collide_cumulant_d3q19(f, tau=0.55, C_w=0.5, solid_mask=torch.zeros((4, 5, 5), dtype=torch.bool,
                                                                                 device=DEV))


# ---------------------------------------------------------------------------
# D3Q27 Geier and legacy cumulant (test_cumulant_d3q27_geier.py,
# test_d3q27_cumulant_reference_invariants.py)
# ---------------------------------------------------------------------------


def _eq27_64(seed, amp, shape=(2, 3, 4)):
    g = torch.Generator().manual_seed(seed)
    rho = 0.9 + 0.2 * torch.rand(shape, generator=g, dtype=torch.float64)
    u = [amp * (2.0 * torch.rand(shape, generator=g, dtype=torch.float64) - 1.0) for _ in range(3)]
    return tl.equilibrium27(rho.to(DEV), *(v.to(DEV) for v in u))


def _perturbed27(seed, amp=0.12, scale=1e-3):
    feq = _eq27_64(seed, amp)
    g = torch.Generator().manual_seed(seed + 1)
    return feq + scale * torch.randn(feq.shape, generator=g, dtype=torch.float64).to(DEV)


@pytest.mark.parametrize("omega_even", [0.4, 1.0, 1.6])
@pytest.mark.parametrize("tau", [0.55, 0.83, 1.7])
def test_geier_fixed_point_and_conservation(tau, omega_even):
    feq = _eq27_64(20260811, 0.12)
    out = tl.collide_cumulant_geier_d3q27(feq, tau=tau, omega_b=1.3, omega_odd=0.7, omega_even=omega_even)
    assert torch.allclose(out, feq, rtol=0.0, atol=1e-12)
    f = _perturbed27(20260811)
    after = tl.collide_cumulant_geier_d3q27(f, tau=tau, omega_b=1.2, omega_odd=1.1, omega_even=0.9)
    assert torch.allclose(_raw_moments(after), _raw_moments(f), rtol=0.0, atol=1e-12)


def test_geier_finite_and_damps_non_equilibrium():
    f = _perturbed27(20260811)
    out = tl.collide_cumulant_geier_d3q27(f, tau=0.5005)
    assert out.shape == f.shape and bool(torch.isfinite(out).all())

    def neq(x):
        return (x - tl.equilibrium27(*tl.macroscopic27(x))).abs().max()

    assert neq(tl.collide_cumulant_geier_d3q27(f, tau=0.55)) < neq(f)


@pytest.mark.parametrize("C_s", [0.0, 0.12])
@pytest.mark.parametrize("tau", [0.55, 0.83, 1.7])
def test_legacy_cumulant27_invariants(tau, C_s):
    before = _perturbed27(20260711, amp=0.04)
    after = tl.collide_cumulant_d3q27(before, tau=tau, omega_b=1.13, omega_odd=0.71, omega_even=1.37, C_s=C_s)
    torch.testing.assert_close(_raw_moments(after), _raw_moments(before), rtol=2e-6, atol=2e-7)
    assert float((_raw_moments(after) - _raw_moments(before)).abs().max()) < 1e-13  # fp64 arithmetic
    feq = _eq27_64(20260712, 0.06)
    out = tl.collide_cumulant_d3q27(feq, tau=tau, omega_b=1.13, omega_odd=0.71, omega_even=1.37, C_s=C_s)
    torch.testing.assert_close(out, feq, rtol=2e-6, atol=2e-7)


# ---------------------------------------------------------------------------
# mass correction (test_correct_mass.py)
# ---------------------------------------------------------------------------


@pytest.mark.parametrize("q,shape,rho0,target", [(9, (10, 12), 1.5, 100.0), (9, (5, 5), 4.0, 50.0),
                                                 (9, (5, 5), 0.5, 100.0), (19, (8, 10, 12), 1.2, 500.0),
                                                 (19, (4, 5, 6), 5.0, 50.0), (27, (6, 8, 10), 1.5, 300.0),
                                                 (27, (3, 4, 5), 6.0, 40.0)])
def test_correct_mass_restores_target(q, shape, rho0, target):
    rho = _full(shape, rho0)
    vel = [torch.zeros_like(rho) for _ in range(2 if q == 9 else 3)]
    eq, cm = {9: (tl.equilibrium, tl.correct_mass), 19: (tl.equilibrium3d, tl.correct_mass3d),
              27: (tl.equilibrium27, tl.correct_mass27)}[q]
    f = eq(rho, *vel)
    out = cm(f, target)
    assert out.shape == f.shape
    mass = float(out.double().sum())
    assert abs(mass - target) < 1e-4
    assert (mass < float(f.double().sum())) == (target < float(f.double().sum()))


@pytest.mark.parametrize("q,shape,slot", [(9, (5, 5), 4), (19, (5, 5, 5), 9), (27, (4, 4, 4), 13)])
def test_correct_mass_near_zero_is_unchanged(q, shape, slot):
    f = torch.zeros((q, *shape), device=DEV)
    f[(slot,) + tuple(s // 2 for s in shape)] = 1e-40
    cm = {9: tl.correct_mass, 19: tl.correct_mass3d, 27: tl.correct_mass27}[q]
    assert torch.allclose(cm(f, 1.0), f)


def test_correct_mass_preserves_velocity():
    shape = (10, 12)
    rho = _full(shape, 2.0)
    f = tl.equilibrium(rho, _full(shape, 0.1), _full(shape, 0.05))
    _, ux0, uy0 = tl.macroscopic(f)
    _, ux1, uy1 = tl.macroscopic(tl.correct_mass(f, 200.0))
    assert float((ux1 - ux0).abs().max()) < 1e-7 and float((uy1 - uy0).abs().max()) < 1e-7
