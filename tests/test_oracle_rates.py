"""Pins of the oracle's relaxation-rate placement and of its gradient closures (no GPU).

MRT and the Geier cumulant collision relax each moment at a named rate.  The
checks here build every moment from its defining polynomial by brute force
over the velocity set (not from the oracle's matrices or rate vectors), give
every rate a distinct value, and require the post-collision moment to satisfy
``m* − m_eq = (1 − s)(m − m_eq)`` with the rate the reference names for it.
A swap of two rates, a rate on the wrong row or a wrong row fails one of them.

The reference's own outputs (``tests/golden/reference_vectors.npz``, written
by ``tests/golden/make_reference_vectors.py``) pin what no polynomial fixes:
the legacy D3Q27 operator, D3Q27 MRT's substituted basis rows, and the
WALE / Vreman closures of the D3Q19 cumulant collision.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import ref
from oracle.lattice import D2Q9, D3Q19, D3Q27
from tests import _fields as F

VECTORS = Path(__file__).parent / "golden" / "reference_vectors.npz"

# distinct rates; 1/tau is the shear rate
TAU = 0.8
RATES = {"s_e": 1.31, "s_eps": 0.47, "s_q": 1.73, "s_pi": 0.91, "s_nu": 1.0 / TAU, "one": 1.0}


def _moment(f, lat, poly):
    c = lat.c.astype(np.float64)
    w = np.array([poly(cx, cy, cz) for cx, cy, cz in c])
    return np.tensordot(w, f, axes=(0, 0))


def _check_rates(lat, f, out, table):
    rho, ux, uy, uz = ref.macroscopic(f)
    feq = ref.equilibrium(lat, rho, ux, uy, uz)
    for name, poly, rate in table:
        m, ms, meq = (_moment(x, lat, poly) for x in (f, out, feq))
        s = RATES[rate]
        assert np.abs(m - meq).max() > 1e-6, name  # the state excites this moment
        np.testing.assert_allclose(ms - meq, (1.0 - s) * (m - meq), rtol=0, atol=1e-13, err_msg=name)


def _e2(cx, cy, cz):
    return cx * cx + cy * cy + cz * cz


# D2Q9 (solver.py:20-30 rows, S vector solver.py:307-322): Lallemand & Luo's basis
D2Q9_MOMENTS = [
    ("e", lambda x, y, z: 3 * _e2(x, y, z) - 4, "s_e"),
    ("eps", lambda x, y, z: 4 - 10.5 * _e2(x, y, z) + 4.5 * _e2(x, y, z) ** 2, "s_eps"),
    ("q_x", lambda x, y, z: (3 * _e2(x, y, z) - 5) * x, "s_q"),
    ("q_y", lambda x, y, z: (3 * _e2(x, y, z) - 5) * y, "s_q"),
    ("p_xx", lambda x, y, z: x * x - y * y, "s_nu"),
    ("p_xy", lambda x, y, z: x * y, "s_nu"),
]

# D3Q19 (rows solver3d.py:24-58, S vector solver3d.py:64-90)
D3Q19_MOMENTS = [
    ("e", lambda x, y, z: 19 * _e2(x, y, z) - 30, "s_e"),
    ("eps", lambda x, y, z: (21 * _e2(x, y, z) ** 2 - 53 * _e2(x, y, z) + 24) / 2, "s_eps"),
    ("q_x", lambda x, y, z: x * _e2(x, y, z) * (5 * _e2(x, y, z) - 9) / 2, "s_q"),
    ("q_y", lambda x, y, z: y * _e2(x, y, z) * (5 * _e2(x, y, z) - 9) / 2, "s_q"),
    ("q_z", lambda x, y, z: z * _e2(x, y, z) * (5 * _e2(x, y, z) - 9) / 2, "s_q"),
    ("3p_xx", lambda x, y, z: 3 * x * x - _e2(x, y, z), "s_nu"),
    ("p_ww", lambda x, y, z: x * x - y * y, "s_nu"),
    ("p_xy", lambda x, y, z: x * y, "s_nu"),
    ("p_xz", lambda x, y, z: x * z, "s_nu"),
    ("p_yz", lambda x, y, z: y * z, "s_nu"),
    ("pi_xx", lambda x, y, z: (3 * _e2(x, y, z) - 5) * (3 * x * x - _e2(x, y, z)) / 2, "s_pi"),
    ("pi_ww", lambda x, y, z: (3 * _e2(x, y, z) - 5) * (x * x - y * y) / 2, "s_pi"),
    ("x2y", lambda x, y, z: x * x * y, "one"),
    ("x2z", lambda x, y, z: x * x * z, "one"),
    ("y2x", lambda x, y, z: y * y * x, "one"),
]

# D3Q27 (d3q27.py:545-633 basis, d3q27.py:681-712 rates): the rows that Gram-Schmidt keeps as
# polynomials.  Raw rows are orthogonalised in order, so a raw polynomial lies in the span of the
# orthogonal rows up to its own index; with the conserved rows removed these have one rate each.
D3Q27_MOMENTS = [
    ("e", _e2, "s_e"),
    ("3c_xx", lambda x, y, z: 3 * x * x - _e2(x, y, z), "s_nu"),
    ("c_yy-c_zz", lambda x, y, z: y * y - z * z, "s_nu"),
    ("c_xy", lambda x, y, z: x * y, "s_nu"),
    ("c_xz", lambda x, y, z: x * z, "s_nu"),
    ("c_yz", lambda x, y, z: y * z, "s_nu"),
    ("c_x e", lambda x, y, z: x * _e2(x, y, z), "s_q"),
    ("c_y e", lambda x, y, z: y * _e2(x, y, z), "s_q"),
    ("c_z e", lambda x, y, z: z * _e2(x, y, z), "s_q"),
    ("c_xxy", lambda x, y, z: x * x * y, "s_q"),
    ("c_xxz", lambda x, y, z: x * x * z, "s_q"),
    ("c_yyx", lambda x, y, z: y * y * x, "s_q"),
    ("c_xyz", lambda x, y, z: x * y * z, "s_pi"),
]


@pytest.mark.parametrize("lat,table", [(D2Q9, D2Q9_MOMENTS), (D3Q19, D3Q19_MOMENTS), (D3Q27, D3Q27_MOMENTS)],
                         ids=["D2Q9", "D3Q19", "D3Q27"])
def test_mrt_moments_relax_at_their_named_rate(lat, table):
    shape = (3, 4, 5) if lat.dim == 3 else (1, 4, 5)
    f = F.random_state(lat.q, shape, seed=31, noise=5e-2)
    if lat.dim == 2:
        f = f[:, 0]
    out = ref.collide_mrt(f, TAU, s_e=RATES["s_e"], s_eps=RATES["s_eps"], s_q=RATES["s_q"],
                          **({} if lat.dim == 2 else {"s_pi": RATES["s_pi"]}))
    _check_rates(lat, f, out, table)


@pytest.mark.parametrize("lat,table", [(D2Q9, D2Q9_MOMENTS), (D3Q19, D3Q19_MOMENTS)], ids=["D2Q9", "D3Q19"])
def test_mrt_polynomials_span_every_non_conserved_direction(lat, table):
    """With the conserved moments the polynomials are a basis: every row of M is checked."""
    c = lat.c.astype(np.float64)
    rows = [np.ones(lat.q), c[:, 0], c[:, 1]] + ([c[:, 2]] if lat.dim == 3 else [])
    rows += [np.array([p(*v) for v in c]) for _, p, _ in table]
    assert np.linalg.matrix_rank(np.array(rows)) == lat.q


def test_mrt27_every_row_against_the_reference():
    """All 27 rows, the substituted ones included, through the reference's own outputs (fp32)."""
    d = np.load(VECTORS)
    for i in range(3):
        f, p, want = d[f"mrt27_{i}_in"], d[f"mrt27_{i}_params"], d[f"mrt27_{i}_out"]
        s_pi = None if np.isnan(p[4]) else p[4]
        got = ref.collide_mrt(f.astype(np.float64), p[0], p[1], p[2], p[3], s_pi)
        np.testing.assert_allclose(got, want, rtol=0, atol=5e-7)
        # the fixture resolves the rate placement: a swap of two rates moves the result by far more
        swapped = ref.collide_mrt(f.astype(np.float64), p[0], p[1], p[3], p[2], s_pi)
        assert np.abs(swapped - want).max() > 1e-5


# ---------------------------------------------------------------------------
# Geier cumulants, per order (cumulant.py:1131-1148)
# ---------------------------------------------------------------------------


def _cumulants(f, u):
    """Central moments by brute force over D3Q27, then the partition formula (already pinned)."""
    c = D3Q27.c.astype(np.float64)
    k = []
    for a, b, g in ref.DEGREES_27:
        w = ((c[:, 0, None, None, None] - u[0]) ** a * (c[:, 1, None, None, None] - u[1]) ** b
             * (c[:, 2, None, None, None] - u[2]) ** g)
        k.append((w * f).sum(0))
    return ref.central_to_cumulant(k)


def test_geier_cumulants_relax_at_their_order_rate():
    f = F.random_state(27, (2, 3, 4), seed=41, noise=5e-2)
    om, ob, oo, oe = 1.0 / 0.7, 1.37, 0.61, 1.83
    out = ref.collide_cumulant_geier_d3q27(f, 0.7, omega_b=ob, omega_odd=oo, omega_even=oe)
    rho, ux, uy, uz = ref.macroscopic(f)
    u = (ux, uy, uz)
    feq = ref.equilibrium(D3Q27, rho, ux, uy, uz)
    c, cs, ce = _cumulants(f, u), _cumulants(out, u), _cumulants(feq, u)
    deg = ref.DEGREES_27
    for i, d in enumerate(deg):
        order = sum(d)
        if order < 2:
            continue
        if order == 2 and max(d) == 2:
            continue  # diagonal: trace and deviators below
        rate = om if order == 2 else (oo if order == 3 else oe)
        assert np.abs(c[i] - ce[i]).max() > 1e-7, d
        np.testing.assert_allclose(cs[i] - ce[i], (1 - rate) * (c[i] - ce[i]), rtol=0, atol=1e-13, err_msg=str(d))
    ixx, iyy, izz = (deg.index(t) for t in ((2, 0, 0), (0, 2, 0), (0, 0, 2)))

    def tr(v):
        return v[ixx] + v[iyy] + v[izz]

    np.testing.assert_allclose(tr(cs) - tr(ce), (1 - ob) * (tr(c) - tr(ce)), rtol=0, atol=1e-13)
    for a, b in ((ixx, iyy), (ixx, izz)):
        np.testing.assert_allclose(cs[a] - cs[b] - (ce[a] - ce[b]), (1 - om) * (c[a] - c[b] - (ce[a] - ce[b])),
                                   rtol=0, atol=1e-13)


# ---------------------------------------------------------------------------
# the legacy D3Q27 operator and the gradient closures against the reference's outputs
# ---------------------------------------------------------------------------


def test_legacy_d3q27_against_the_reference():
    d = np.load(VECTORS)
    for i in range(5):
        f, p, want = d[f"legacy27_{i}_in"], d[f"legacy27_{i}_params"], d[f"legacy27_{i}_out"]
        got = ref.collide_cumulant_d3q27(f, *p)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-14, err_msg=f"case {i}")


def _sgs_args(d, i):
    p = d[f"sgs19_{i}_params"]
    mask = d[f"sgs19_{i}_mask"] if bool(d[f"sgs19_{i}_has_mask"]) else None
    return dict(tau=p[0], omega_b=p[1], omega_even=p[2], C_w=p[3], C_v=p[4], solid_mask=mask,
                wall_velocity=tuple(p[5:8]))


def test_wale_vreman_cumulant_d3q19_against_the_reference():
    d = np.load(VECTORS)
    for i in range(6):
        f, want = d[f"sgs19_{i}_in"], d[f"sgs19_{i}_out"]
        got = ref.collide_cumulant_d3q19(f, **_sgs_args(d, i))
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-14, err_msg=f"case {i}")
        # the closure moves the result well beyond that tolerance
        a = _sgs_args(d, i)
        plain = ref.collide_cumulant_d3q19(f, a["tau"], a["omega_b"], 1.0, a["omega_even"])
        assert np.abs(plain - want).max() > 1e-5


def _grid(shape):
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in shape), indexing="ij")
    return z, y, x


def test_edge_derivative_exact_on_quadratics_and_short_axes():
    z, y, x = _grid((4, 5, 6))
    field = 0.3 * x * x - 0.7 * x * y + 0.2 * z * z + 1.5 * y
    np.testing.assert_allclose(ref.edge_derivative(field, 2), 0.6 * x - 0.7 * y, atol=1e-13)
    np.testing.assert_allclose(ref.edge_derivative(field, 1), -0.7 * x + 1.5, atol=1e-13)
    np.testing.assert_allclose(ref.edge_derivative(field, 0), 0.4 * z, atol=1e-13)
    g = np.arange(10.0).reshape(2, 5) ** 2
    np.testing.assert_allclose(ref.edge_derivative(g, 0), np.stack([g[1] - g[0]] * 2))
    assert not ref.edge_derivative(g[:1], 0).any()


def test_wale_closed_forms():
    z, y, x = _grid((5, 6, 7))
    zero = np.zeros_like(x)
    # pure shear: g·g = 0, so WALE adds nothing
    assert np.abs(ref.wale_nu_t(0.01 * y, zero, zero, 0.5)).max() < 1e-20
    # rigid rotation at rate W: |S| = 0, |S^d|^2 = 2 W^4 / 3, nu_t = C_w^2 (2/3)^(1/4) W
    w = 0.02
    nu = ref.wale_nu_t(-w * y, w * x, zero, 0.5)
    np.testing.assert_allclose(nu, 0.25 * (2.0 / 3.0) ** 0.25 * w, rtol=1e-10)


def test_vreman_closed_forms():
    z, y, x = _grid((5, 6, 7))
    zero = np.zeros_like(x)
    # a velocity that varies in one direction only: beta has rank one, B_beta = 0
    assert np.abs(ref.vreman_nu_t(0.01 * y + 0.002 * y * y, 0.03 * y, zero, 0.07)).max() < 1e-12
    # plane strain u = (a x, -a y, 0): B = a^4, |alpha|^2 = 2 a^2, nu_t = C_v |a| / sqrt(2)
    a = 0.013
    np.testing.assert_allclose(ref.vreman_nu_t(a * x, -a * y, zero, 0.07), 0.07 * a / np.sqrt(2.0), rtol=1e-10)


def test_gradient_closure_uses_wall_velocity_on_solid_cells():
    f = F.random_state(19, (3, 4, 5), seed=7, noise=1e-3)
    mask = np.zeros((3, 4, 5), bool)
    mask[:, :, 0] = True
    rho, ux, uy, uz = ref.macroscopic(f)
    ux, uy, uz = np.where(mask, 0.04, ux), np.where(mask, 0.0, uy), np.where(mask, 0.0, uz)
    want = np.clip(0.6 + 3.0 * ref.wale_nu_t(ux, uy, uz, 0.5), ref.TAU_EFF_MIN, ref.tau_eff_max(0.6))
    got = ref.gradient_closure_tau(f, 0.6, C_w=0.5, solid_mask=mask, wall_velocity=(0.04, 0.0, 0.0))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-15)
