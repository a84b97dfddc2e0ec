"""numpy restatement of the reference per-step lattice update (test infrastructure).

Populations are numpy arrays ``f[Q, nz, ny, nx]`` (x fastest, SoA, the
reference layout, ``d3q19.py:95-103``); D2Q9 fields ``f[9, ny, nx]`` are
handled as ``nz = 1``.  Arithmetic follows the input dtype (float32 or
float64), with python-float scalars, so a float32 input reproduces the
reference's float32 evaluation up to summation order.

Every function cites the reference file:line it restates.  How each one is
pinned is recorded in ``tests/test_oracle_*.py``.
"""

from __future__ import annotations

from fractions import Fraction

import numpy as np

from .lattice import D2Q9, D3Q19, D3Q27, Lattice, for_q

CS2 = 1.0 / 3.0
RHO_FLOOR = 1e-12
TAU_EFF_MIN = 0.5001  # turbulence.py:77


# ---------------------------------------------------------------------------
# helpers
# ---------------------------------------------------------------------------


def _lat(f: np.ndarray) -> Lattice:
    return for_q(f.shape[0])


def _to4(f: np.ndarray) -> tuple[np.ndarray, bool]:
    if f.ndim == 3:
        return f[:, None], True
    if f.ndim == 4:
        return f, False
    raise ValueError(f"populations must be (Q,ny,nx) or (Q,nz,ny,nx), got {f.shape}")


def _from4(f4: np.ndarray, was2d: bool) -> np.ndarray:
    return f4[:, 0] if was2d else f4


def _cvec(lat: Lattice, axis: int, dtype, ndim: int) -> np.ndarray:
    return lat.c[:, axis].astype(dtype).reshape((lat.q,) + (1,) * ndim)


# ---------------------------------------------------------------------------
# equilibrium / macroscopic
# ---------------------------------------------------------------------------


def equilibrium(lat: Lattice, rho, ux, uy, uz=None) -> np.ndarray:
    """Second-order Hermite equilibrium ``w ρ (1 + 3cu + 4.5cu² − 1.5u²)``.

    d2q9.py:39-58 (D2Q9), d3q19.py:83-125 (D3Q19), d3q27.py:353-392 (D3Q27).
    Weights are the exact rationals cast to the field dtype
    (``W_EXACT64`` convention, d3q19.py:61, d3q27.py:66).
    """
    rho = np.asarray(rho)
    if uz is None:
        uz = np.zeros_like(rho)
    if not (rho.shape == np.shape(ux) == np.shape(uy) == np.shape(uz)):
        raise ValueError("rho, ux, uy, and uz shapes must match")
    dt = rho.dtype
    nd = rho.ndim
    w = lat.w.astype(dt).reshape((lat.q,) + (1,) * nd)
    cx, cy, cz = (_cvec(lat, a, dt, nd) for a in range(3))
    u_sq = ux * ux + uy * uy + uz * uz
    cu = cx * ux + cy * uy + cz * uz
    return w * rho[None] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * u_sq[None])


def macroscopic(f: np.ndarray):
    """``ρ = Σ f``, ``u = Σ c f / max(ρ, 1e-12)``.

    d2q9.py:61-74, d3q19.py:165-186, d3q27.py:395-420.  Returns
    ``(rho, ux, uy, uz)``; ``uz`` is zero for D2Q9.
    """
    lat = _lat(f)
    nd = f.ndim - 1
    dt = f.dtype
    rho = f.sum(axis=0)
    rho_safe = np.maximum(rho, dt.type(RHO_FLOOR))
    u = [(f * _cvec(lat, a, dt, nd)).sum(axis=0) / rho_safe for a in range(3)]
    return rho, u[0], u[1], u[2]


def neq_stress(f_neq: np.ndarray):
    """Π_ab = Σ_q c_a c_b f_neq (six components, xx yy zz xy xz yz).

    turbulence.py:143-180 (3-D), cumulant.py:1254-1264, triton_fused_obstacle.py:1326-1332.
    """
    lat = _lat(f_neq)
    nd = f_neq.ndim - 1
    c = [_cvec(lat, a, f_neq.dtype, nd) for a in range(3)]
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    return tuple((c[a] * c[b] * f_neq).sum(axis=0) for a, b in pairs)


# ---------------------------------------------------------------------------
# streaming / bounce-back
# ---------------------------------------------------------------------------


def stream(f: np.ndarray) -> np.ndarray:
    """Periodic pull stream ``out[q, x] = f[q, x − c_q]``.

    solver.py:396-417 (D2Q9 gather), solver3d.py:210-239 / 266-284 (D3Q19),
    d3q27.py:766-791 (D3Q27 roll).  A pure permutation, so it is bit-exact.
    """
    lat = _lat(f)
    f4, was2d = _to4(f)
    out = np.empty_like(f4)
    for q in range(lat.q):
        cx, cy, cz = (int(v) for v in lat.c[q])
        out[q] = np.roll(f4[q], shift=(cz, cy, cx), axis=(0, 1, 2))
    return _from4(out, was2d)


def bounce_back(f: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """Full-way swap ``where(mask, f[opp], f)``.

    boundaries.py:39-47, boundaries3d.py:112-132, boundaries_d3q27.py:45-60.
    """
    lat = _lat(f)
    return np.where(mask[None], f[lat.opposite], f)


def fullway_pull(f: np.ndarray, solid: np.ndarray) -> np.ndarray:
    """Pull form of ``bounce_back(stream(f), solid)``.

    Fluid x: ``f[q, x − c_q]``; solid x: ``f[opp(q), x + c_q]``
    (triton_fused_obstacle.py:1168-1222).  Equal to ``bounce_back(stream(f), solid)``.
    """
    return bounce_back(stream(f), solid)


# ---------------------------------------------------------------------------
# BGK
# ---------------------------------------------------------------------------


def collide_bgk(f: np.ndarray, tau: float) -> np.ndarray:
    """``f − (f − f_eq)/τ``: solver.py:67-71, solver3d.py:100-104, d3q27.py:423-435."""
    lat = _lat(f)
    rho, ux, uy, uz = macroscopic(f)
    feq = equilibrium(lat, rho, ux, uy, uz)
    return f - (f - feq) / tau


def _feq_safe(f: np.ndarray):
    """Fused-kernel moments: equilibrium built with the floored density.

    triton_fused.py:419-427 and triton_fused_obstacle.py:1262-1272 use
    ``rho_safe`` inside f_eq (the eager ops use ``rho``); identical unless
    ρ < 1e-12.
    """
    lat = _lat(f)
    rho, ux, uy, uz = macroscopic(f)
    rho_safe = np.maximum(rho, f.dtype.type(RHO_FLOOR))
    return equilibrium(lat, rho_safe, ux, uy, uz), rho_safe, (ux, uy, uz)


def fused_step(f: np.ndarray, tau: float) -> np.ndarray:
    """Periodic pull-stream then BGK with ``ω = 1/τ`` (triton_fused.py:396-436)."""
    fin = stream(f)
    feq, _, _ = _feq_safe(fin)
    return fin - (1.0 / tau) * (fin - feq)


# ---------------------------------------------------------------------------
# Smagorinsky closure
# ---------------------------------------------------------------------------


def tau_eff_max(tau: float) -> float:
    """turbulence.py:93-100: ``max(1, τ + 0.5)``."""
    return max(1.0, tau + 0.5)


def smagorinsky_tau(tau: float, pi_norm, rho, cs_delta_sq: float):
    """Hou et al. self-consistent τ_eff, clamped (turbulence.py:183-208).

    ``τ_eff = ½(τ + √(τ² + 18 Cs²Δ² |Π|/ρ))`` clamped to
    ``[0.5001, max(1, τ + 0.5)]``; kernel form triton_fused_obstacle.py:295-335.
    """
    dt = np.asarray(pi_norm).dtype
    rho_safe = np.maximum(rho, dt.type(RHO_FLOOR))
    disc = tau * tau + 18.0 * cs_delta_sq * pi_norm / rho_safe
    te = 0.5 * (tau + np.sqrt(np.maximum(disc, dt.type(0.0))))
    return np.clip(te, TAU_EFF_MIN, tau_eff_max(tau)).astype(dt)


def pi_norm(f_neq: np.ndarray, floor: float = 1e-30):
    """Frobenius norm ``√(Σ Π_aa² + 2Σ Π_ab²)`` with the 1e-30 radicand floor (turbulence.py:143-180)."""
    pxx, pyy, pzz, pxy, pxz, pyz = neq_stress(f_neq)
    rad = pxx**2 + pyy**2 + pzz**2 + 2.0 * (pxy**2 + pxz**2 + pyz**2)
    return np.sqrt(np.maximum(rad, f_neq.dtype.type(floor)))


def collide_smagorinsky_bgk(f: np.ndarray, tau: float, C_s: float = 0.1) -> np.ndarray:
    """BGK with per-cell Smagorinsky τ_eff (turbulence.py:216-267, 472-494)."""
    lat = _lat(f)
    rho, ux, uy, uz = macroscopic(f)
    feq = equilibrium(lat, rho, ux, uy, uz)
    fneq = f - feq
    te = smagorinsky_tau(tau, pi_norm(fneq), rho, C_s * C_s)
    return f - fneq / te[None]


# ---------------------------------------------------------------------------
# MRT
# ---------------------------------------------------------------------------

_M_D2Q9 = [
    [1, 1, 1, 1, 1, 1, 1, 1, 1],
    [-4, -1, -1, -1, -1, 2, 2, 2, 2],
    [4, -2, -2, -2, -2, 1, 1, 1, 1],
    [0, 1, 0, -1, 0, 1, -1, -1, 1],
    [0, -2, 0, 2, 0, 1, -1, -1, 1],
    [0, 0, 1, 0, -1, 1, 1, -1, -1],
    [0, 0, -2, 0, 2, 1, 1, -1, -1],
    [0, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 1, -1, 1, -1],
]  # solver.py:20-30


def _frac_inverse(m: list[list[Fraction]]) -> list[list[Fraction]]:
    """Exact Gauss-Jordan inverse over the rationals."""
    n = len(m)
    a = [list(row) + [Fraction(int(i == j)) for j in range(n)] for i, row in enumerate(m)]
    for col in range(n):
        piv = next(r for r in range(col, n) if a[r][col] != 0)
        a[col], a[piv] = a[piv], a[col]
        p = a[col][col]
        a[col] = [v / p for v in a[col]]
        for r in range(n):
            if r != col and a[r][col] != 0:
                fac = a[r][col]
                a[r] = [vr - fac * vc for vr, vc in zip(a[r], a[col])]
    return [row[n:] for row in a]


def _mrt_rows_d3q19() -> list[list[Fraction]]:
    """D3Q19 moment rows (solver3d.py:24-58)."""
    rows = []
    for cx, cy, cz in D3Q19.c.tolist():
        e2 = cx * cx + cy * cy + cz * cz
        e4 = e2 * e2
        rows.append(
            [
                Fraction(1),
                Fraction(19 * e2 - 30),
                Fraction(21 * e4 - 53 * e2 + 24, 2),
                Fraction(cx),
                Fraction(cx * e2 * (5 * e2 - 9), 2),
                Fraction(cy),
                Fraction(cy * e2 * (5 * e2 - 9), 2),
                Fraction(cz),
                Fraction(cz * e2 * (5 * e2 - 9), 2),
                Fraction(3 * cx * cx - e2),
                Fraction(cx * cx - cy * cy),
                Fraction(cx * cy),
                Fraction(cx * cz),
                Fraction(cy * cz),
                Fraction((3 * e2 - 5) * (3 * cx * cx - e2), 2),
                Fraction((3 * e2 - 5) * (cx * cx - cy * cy), 2),
                Fraction(cx * cx * cy),
                Fraction(cx * cx * cz),
                Fraction(cy * cy * cx),
            ]
        )
    return [list(r) for r in zip(*rows)]  # transpose: rows are moments


def _mrt_rows_d3q27() -> list[list[Fraction]]:
    """D3Q27 Gram-Schmidt basis (d3q27.py:545-633), exact in rationals.

    Raw rows are orthogonalised in order against the rows kept so far;
    a row that becomes zero is replaced by the first standard basis vector
    whose orthogonal remainder is non-zero.
    """
    raw = []
    for cx, cy, cz in D3Q27.c.tolist():
        e2 = cx * cx + cy * cy + cz * cz
        raw.append(
            [
                1, cx, cy, cz, e2, 3 * cx * cx - e2, cy * cy - cz * cz, cx * cy, cx * cz, cy * cz,
                cx * e2, cy * e2, cz * e2, cx * cx * cy, cx * cx * cz, cy * cy * cx, cy * cy * cz,
                cz * cz * cx, cz * cz * cy, e2 * e2, cx * cx * e2, cy * cy * e2, cz * cz * e2,
                cx * cx * cy * cy, cx * cx * cz * cz, cy * cy * cz * cz, cx * cy * cz,
            ]
        )
    raw_rows = [[Fraction(v) for v in r] for r in zip(*raw)]

    def dot(a, b):
        return sum(x * y for x, y in zip(a, b))

    orth: list[list[Fraction]] = []
    for row in raw_rows:
        v = list(row)
        for p in orth:
            k = dot(v, p) / dot(p, p)
            v = [a - k * b for a, b in zip(v, p)]
        if all(x == 0 for x in v):
            for i in range(27):
                u = [Fraction(int(j == i)) for j in range(27)]
                for p in orth:
                    k = dot(u, p) / dot(p, p)
                    u = [a - k * b for a, b in zip(u, p)]
                if any(x != 0 for x in u):
                    v = u
                    break
        orth.append(v)
    return orth


_MRT_CACHE: dict[str, tuple[np.ndarray, np.ndarray]] = {}


def mrt_matrices(lat: Lattice) -> tuple[np.ndarray, np.ndarray]:
    """Exact (M, M⁻¹) in binary64 for the three lattices."""
    if lat.name not in _MRT_CACHE:
        if lat.name == "D2Q9":
            m = [[Fraction(v) for v in r] for r in _M_D2Q9]
        elif lat.name == "D3Q19":
            m = _mrt_rows_d3q19()
        else:
            m = _mrt_rows_d3q27()
        mi = _frac_inverse(m)
        _MRT_CACHE[lat.name] = (
            np.array([[float(v) for v in r] for r in m]),
            np.array([[float(v) for v in r] for r in mi]),
        )
    return _MRT_CACHE[lat.name]


def mrt_rates(lat: Lattice, tau: float, s_e: float, s_eps: float, s_q: float, s_pi=None):
    """Relaxation vector S (solver.py:307-322, solver3d.py:61-90, d3q27.py:681-712)."""
    s_nu = 1.0 / tau
    if lat.name == "D2Q9":
        return [0.0, s_e, s_eps, 0.0, s_q, 0.0, s_q, s_nu, s_nu]
    if s_pi is None:
        s_pi = s_e
    if lat.name == "D3Q19":
        return [0.0, s_e, s_eps, 0.0, s_q, 0.0, s_q, 0.0, s_q] + [s_nu] * 5 + [s_pi, s_pi, 1.0, 1.0, 1.0]
    return [0.0] * 4 + [s_e] + [s_nu] * 5 + [s_q] * 9 + [s_eps] + [s_pi] * 7


def collide_smagorinsky_mrt(f: np.ndarray, tau: float, C_s: float = 0.1, s_e=None, s_eps=None, s_q=None,
                            s_pi=None) -> np.ndarray:
    """MRT with the stress rows relaxed at the per-cell Smagorinsky 1/τ_eff
    (turbulence.py:270-342 D2Q9, :345-433 D3Q19): τ_eff from |Π^neq| of f − f_eq(ρ, u)."""
    lat = _lat(f)
    rho, ux, uy, uz = macroscopic(f)
    te = smagorinsky_tau(tau, pi_norm(f - equilibrium(lat, rho, ux, uy, uz)), rho, C_s * C_s)
    return collide_mrt(f, tau, s_e, s_eps, s_q, s_pi, tau_cell=te)


def collide_mrt(f: np.ndarray, tau: float, s_e=None, s_eps=None, s_q=None, s_pi=None, *,
                tau_cell=None) -> np.ndarray:
    """``M⁻¹[m − S(m − m_eq)]`` (solver.py:249-335, solver3d.py:107-160, d3q27.py:643-726).

    Defaults: D2Q9 ``s_e=1.64, s_eps=1.54, s_q=1.7``; D3Q19/D3Q27
    ``s_e=1.19, s_eps=1.4, s_q=1.2``, ``s_pi=s_e``.
    """
    lat = _lat(f)
    d2 = lat.name == "D2Q9"
    s_e = (1.64 if d2 else 1.19) if s_e is None else s_e
    s_eps = (1.54 if d2 else 1.4) if s_eps is None else s_eps
    s_q = (1.7 if d2 else 1.2) if s_q is None else s_q
    m, mi = mrt_matrices(lat)
    dt = f.dtype
    m, mi = m.astype(dt), mi.astype(dt)
    s = np.array(mrt_rates(lat, tau, s_e, s_eps, s_q, s_pi), dtype=dt)
    shape = f.shape
    ff = f.reshape(lat.q, -1)
    rho, ux, uy, uz = macroscopic(f)
    feq = equilibrium(lat, rho, ux, uy, uz).reshape(lat.q, -1)
    mom = m @ ff
    mom_eq = m @ feq
    rates = np.broadcast_to(s[:, None], mom.shape).copy()
    if tau_cell is not None:  # the 1/τ rows take the per-cell rate
        shear = np.array(mrt_rates(lat, -1.0, 0.0, 0.0, 0.0, 0.0)) == -1.0
        rates[shear] = 1.0 / np.asarray(tau_cell, dtype=dt).reshape(1, -1)
    return (mi @ (mom - rates * (mom - mom_eq))).reshape(shape)


# ---------------------------------------------------------------------------
# Cascaded central moments (D3Q19 / D3Q27)
# ---------------------------------------------------------------------------

DEGREES_19 = [
    (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2),
    (1, 1, 0), (1, 0, 1), (0, 1, 1), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1),
    (1, 0, 2), (0, 1, 2), (2, 2, 0), (2, 0, 2), (0, 2, 2),
]  # cascaded_collision.py:60-80
DEGREES_27 = [
    (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2),
    (1, 1, 0), (1, 0, 1), (0, 1, 1), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1),
    (1, 0, 2), (0, 1, 2), (1, 1, 1), (2, 2, 0), (2, 0, 2), (0, 2, 2), (2, 1, 1),
    (1, 2, 1), (1, 1, 2), (2, 2, 1), (2, 1, 2), (1, 2, 2), (2, 2, 2),
]  # cascaded_collision.py:82-110, cumulant.py:319-347


def shift_groups(degrees):
    """Triplets (i0, i1, i2) with degree 0/1/2 in one axis (cascaded_collision.py:134-162)."""
    groups = []
    for dim in range(3):
        other = [d for d in range(3) if d != dim]
        bucket: dict = {}
        for idx, deg in enumerate(degrees):
            bucket.setdefault((deg[other[0]], deg[other[1]]), {})[deg[dim]] = idx
        g = []
        for key in sorted(bucket):
            mm = bucket[key]
            if 0 in mm and 1 in mm and 2 in mm:
                g.append((mm[0], mm[1], mm[2]))
        groups.append(g)
    return groups


def moment_matrix(lat: Lattice, degrees) -> tuple[np.ndarray, np.ndarray]:
    """Raw-moment matrix ``M[i, j] = c_jx^a c_jy^b c_jz^c`` and its exact inverse."""
    rows = [
        [Fraction(int(cx) ** a * int(cy) ** b * int(cz) ** d) for cx, cy, cz in lat.c.tolist()]
        for a, b, d in degrees
    ]
    inv = _frac_inverse(rows)
    return (
        np.array([[float(v) for v in r] for r in rows]),
        np.array([[float(v) for v in r] for r in inv]),
    )


def _shift(m: list, groups, u, sign: float) -> list:
    out = list(m)
    for i0, i1, i2 in groups:
        m0, m1, m2 = m[i0], m[i1], m[i2]
        out[i1] = m1 + sign * u * m0
        out[i2] = m2 + sign * 2.0 * u * m1 + u * u * m0
    return out


def to_central(m: list, degrees, ux, uy, uz) -> list:
    """x, then y, then z forward binomial shift (cascaded_collision.py:216-300)."""
    gx, gy, gz = shift_groups(degrees)
    m = _shift(m, gx, ux, -1.0)
    m = _shift(m, gy, uy, -1.0)
    return _shift(m, gz, uz, -1.0)


def to_raw(k: list, degrees, ux, uy, uz) -> list:
    """z, then y, then x inverse shift."""
    gx, gy, gz = shift_groups(degrees)
    k = _shift(k, gz, uz, 1.0)
    k = _shift(k, gy, uy, 1.0)
    return _shift(k, gx, ux, 1.0)


def _relax_second(k: list, omega, omega_bulk) -> None:
    """Trace/deviatoric split at indices 4..9 (cascaded_collision.py:330-384)."""
    tr = k[4] + k[5] + k[6]
    tr_s = (1.0 - omega_bulk) * tr
    for i in (4, 5, 6):
        k[i] = (1.0 - omega) * (k[i] - tr / 3.0) + tr_s / 3.0
    for i in (7, 8, 9):
        k[i] = (1.0 - omega) * k[i]


def collide_cascaded(f: np.ndarray, tau=None, s_bulk=1.0, s_3=1.0, s_4=1.0, s_5=None, s_6=None, *, omega=None):
    """Cascaded central-moment collision on f_neq.

    D3Q19: cascaded_collision.py:445-509; D3Q27: :512-.  ``omega`` may be a
    per-cell array (the fused kernel's CM branch, triton_fused_obstacle.py:1340-1427,
    which is this operator with ``s_bulk = s_3 = s_4 = 1``).
    """
    lat = _lat(f)
    degrees = DEGREES_19 if lat.q == 19 else DEGREES_27
    if omega is None:
        omega = 1.0 / tau
    dt = f.dtype
    m_mat, m_inv = moment_matrix(lat, degrees)
    m_mat, m_inv = m_mat.astype(dt), m_inv.astype(dt)
    rho, ux, uy, uz = macroscopic(f)
    feq = equilibrium(lat, rho, ux, uy, uz)
    fneq = f - feq
    shape = f.shape
    mn = m_mat @ fneq.reshape(lat.q, -1)
    m = [mn[i].reshape(shape[1:]) for i in range(lat.q)]
    k = to_central(m, degrees, ux, uy, uz)
    _relax_second(k, omega, s_bulk)
    if lat.q == 19:
        for i in range(10, 16):
            k[i] = (1.0 - s_3) * k[i]
        for i in range(16, 19):
            k[i] = (1.0 - s_4) * k[i]
    else:
        s_5 = s_4 if s_5 is None else s_5
        s_6 = s_4 if s_6 is None else s_6
        for lo, hi, s in ((10, 17, s_3), (17, 23, s_4), (23, 26, s_5), (26, 27, s_6)):
            for i in range(lo, hi):
                k[i] = (1.0 - s) * k[i]
    ms = to_raw(k, degrees, ux, uy, uz)
    ms = np.stack([np.broadcast_to(v, shape[1:]) for v in ms]).reshape(lat.q, -1)
    return feq + (m_inv @ ms).reshape(shape)


# ---------------------------------------------------------------------------
# D3Q19 Hermite-regularised cumulant
# ---------------------------------------------------------------------------


def hermite_reconstruct(lat: Lattice, pis, dtype) -> np.ndarray:
    """``4.5 w Σ H_ab Π_ab`` with off-diagonal terms doubled (cumulant.py:1331-1362)."""
    nd = np.ndim(pis[0])
    c = [_cvec(lat, a, dtype, nd) for a in range(3)]
    w = lat.w.astype(dtype).reshape((lat.q,) + (1,) * nd)
    pxx, pyy, pzz, pxy, pxz, pyz = pis
    h = (
        (c[0] * c[0] - CS2) * pxx
        + (c[1] * c[1] - CS2) * pyy
        + (c[2] * c[2] - CS2) * pzz
        + 2.0 * (c[0] * c[1]) * pxy
        + 2.0 * (c[0] * c[2]) * pxz
        + 2.0 * (c[1] * c[2]) * pyz
    )
    return 4.5 * w * h


def edge_derivative(field: np.ndarray, axis: int) -> np.ndarray:
    """d/dx along ``axis``: centred inside, second-order one-sided on the two end planes; an axis of
    length 2 gets the forward difference on both planes, of length 1 zero (turbulence.py:595-640)."""
    n = field.shape[axis]
    if n == 1:
        return np.zeros_like(field)
    a = np.moveaxis(field, axis, 0)
    if n == 2:
        d = np.broadcast_to(a[1] - a[0], a.shape).copy()
        return np.moveaxis(d, 0, axis)
    d = np.empty_like(a)
    d[1:-1] = 0.5 * (a[2:] - a[:-2])
    d[0] = 0.5 * (-3.0 * a[0] + 4.0 * a[1] - a[2])
    d[-1] = 0.5 * (3.0 * a[-1] - 4.0 * a[-2] + a[-3])
    return np.moveaxis(d, 0, axis)


def velocity_gradient(ux, uy, uz):
    """``g[i][j] = ∂u_i/∂x_j`` on a (nz, ny, nx) grid, x = axis 2 (turbulence.py:662-695)."""
    return [[edge_derivative(u, ax) for ax in (2, 1, 0)] for u in (ux, uy, uz)]


def wale_nu_t(ux, uy, uz, C_w: float):
    """WALE eddy viscosity ``C_w² |S^d|³ / (|S|⁵ + |S^d|^{5/2} + 1e-30)`` with Δ = 1; S^d the
    traceless symmetric part of g·g (turbulence.py:754-841)."""
    g = velocity_gradient(ux, uy, uz)
    g2 = [[sum(g[i][k] * g[k][j] for k in range(3)) for j in range(3)] for i in range(3)]
    tr = g2[0][0] + g2[1][1] + g2[2][2]
    sd_n2 = sum((g2[i][i] - tr / 3.0) ** 2 for i in range(3))
    sd_n2 = sd_n2 + 2.0 * sum((0.5 * (g2[i][j] + g2[j][i])) ** 2 for i, j in ((0, 1), (0, 2), (1, 2)))
    s_n2 = sum(g[i][i] ** 2 for i in range(3))
    s_n2 = s_n2 + 2.0 * sum((0.5 * (g[i][j] + g[j][i])) ** 2 for i, j in ((0, 1), (0, 2), (1, 2)))
    sd_n2 = np.maximum(sd_n2, 0.0)
    s_n2 = np.maximum(s_n2, 0.0)
    return (C_w * C_w) * sd_n2**1.5 / (s_n2**2.5 + sd_n2**1.25 + 1e-30)


def vreman_nu_t(ux, uy, uz, C_v: float):
    """Vreman (2004) eddy viscosity ``C_v √(max(B_β, 0) / (|α|² + 1e-30))``, β = αᵀα, α_ij = ∂u_i/∂x_j,
    B_β the sum of β's principal 2×2 minors (turbulence.py:844-872)."""
    g = velocity_gradient(ux, uy, uz)
    beta = [[sum(g[k][i] * g[k][j] for k in range(3)) for j in range(3)] for i in range(3)]
    a2 = sum(g[i][j] ** 2 for i in range(3) for j in range(3))
    bb = sum(beta[i][i] * beta[j][j] - beta[i][j] ** 2 for i, j in ((0, 1), (0, 2), (1, 2)))
    return C_v * np.sqrt(np.maximum(bb, 0.0) / (a2 + 1e-30))


def gradient_closure_tau(f, tau: float, C_w: float = 0.0, C_v: float = 0.0, solid_mask=None,
                         wall_velocity=(0.0, 0.0, 0.0)):
    """Per-cell ``τ_eff = clip(τ + 3 ν_t, 0.5001, max(1, τ + 0.5))`` of the WALE or Vreman closure
    (turbulence.py:875-888), from the macroscopic velocity with solid cells set to ``wall_velocity``
    (cumulant.py:1263-1303)."""
    _, ux, uy, uz = macroscopic(f)
    if solid_mask is not None:
        m = np.asarray(solid_mask, bool)
        ux, uy, uz = (np.where(m, f.dtype.type(w), u) for u, w in zip((ux, uy, uz), wall_velocity))
    nu = wale_nu_t(ux, uy, uz, C_w) if C_w > 0.0 else vreman_nu_t(ux, uy, uz, C_v)
    return np.clip(tau + 3.0 * nu, TAU_EFF_MIN, tau_eff_max(tau)).astype(f.dtype)


def collide_cumulant_d3q19(f, tau, omega_b=1.0, omega_odd=1.0, omega_even=1.0, C_s=0.0, C_w=0.0, C_v=0.0,
                           solid_mask=None, wall_velocity=(0.0, 0.0, 0.0), *, omega=None):
    """D3Q19 Hermite-regularised "cumulant" (cumulant.py:1159-1377).

    With ``C_s > 0`` the closure is the UNCLAMPED Hou form of that function
    (cumulant.py:1316-1321); ``C_w`` / ``C_v`` select the WALE / Vreman
    gradient closures (cumulant.py:1263-1303), whose velocity field takes
    ``wall_velocity`` on ``solid_mask``.  ``omega`` overrides with a per-cell
    rate (the fused kernel's CUMULANT branch, triton_fused_obstacle.py:1428-1478).
    ``omega_odd`` is accepted and unused, as in the reference.
    """
    del omega_odd
    lat = _lat(f)
    rho, ux, uy, uz = macroscopic(f)
    feq = equilibrium(lat, rho, ux, uy, uz)
    fneq = f - feq
    pis = neq_stress(fneq)
    if omega is None and (C_w > 0.0 or C_v > 0.0):
        omega = 1.0 / gradient_closure_tau(f, tau, C_w, C_v, solid_mask, wall_velocity)
    if omega is None:
        if C_s > 0.0:
            pxx, pyy, pzz, pxy, pxz, pyz = pis
            pn = np.sqrt(pxx**2 + pyy**2 + pzz**2 + 2.0 * (pxy**2 + pxz**2 + pyz**2))
            rho_safe = np.maximum(rho, f.dtype.type(RHO_FLOOR))
            te = 0.5 * (tau + np.sqrt(tau * tau + 18.0 * C_s * C_s * pn / rho_safe))
            omega = 1.0 / te
        else:
            omega = 1.0 / tau
    pxx, pyy, pzz, pxy, pxz, pyz = pis
    tr = pxx + pyy + pzz
    db = (omega_b - omega) * tr / 3.0
    relaxed = (
        pxx - omega * pxx - db,
        pyy - omega * pyy - db,
        pzz - omega * pzz - db,
        pxy - omega * pxy,
        pxz - omega * pxz,
        pyz - omega * pyz,
    )
    freg = hermite_reconstruct(lat, relaxed, f.dtype)
    fho = fneq - hermite_reconstruct(lat, pis, f.dtype)
    return feq + freg + (1.0 - omega_even) * fho


# ---------------------------------------------------------------------------
# D3Q27 cumulants (Geier 2015): generic partition transform
# ---------------------------------------------------------------------------


def _set_partitions(items):
    if not items:
        yield []
        return
    first, rest = items[0], items[1:]
    for part in _set_partitions(rest):
        for i in range(len(part)):
            yield part[:i] + [[first] + part[i]] + part[i + 1 :]
        yield [[first]] + part


def _block_degree(block):
    d = [0, 0, 0]
    for axis in block:
        d[axis] += 1
    return tuple(d)


def _partitions_for(deg):
    """Set partitions of the multiset of axes of ``deg`` with no singleton blocks."""
    items = [0] * deg[0] + [1] * deg[1] + [2] * deg[2]
    out = []
    for part in _set_partitions(list(range(len(items)))):
        if any(len(b) == 1 for b in part):
            continue
        out.append([_block_degree([items[i] for i in b]) for b in part])
    return out


def central_to_cumulant(k: list) -> list:
    """Normalised central moments → cumulants by the partition formula.

    ``C̃_α = Σ_π (−1)^{|π|−1} (|π|−1)! Π_{B∈π} κ̃_B`` over set partitions
    of the index multiset of α (first central moments vanish, so singleton
    blocks drop out).  ``C = ρ C̃``, ``C[0] = ρ``.  This is the definition
    behind cumulant.py:674-744 derived independently, not restated line by line.
    """
    rho = k[0]
    idx = {d: i for i, d in enumerate(DEGREES_27)}
    kt = [v / rho for v in k]
    out = list(k)
    for i, deg in enumerate(DEGREES_27):
        if sum(deg) < 2:
            continue
        acc = 0.0
        for part in _partitions_for(deg):
            n = len(part)
            term = ((-1.0) ** (n - 1)) * float(_factorial(n - 1))
            prod = 1.0
            for b in part:
                prod = prod * kt[idx[b]]
            acc = acc + term * prod
        out[i] = acc * rho
    out[0] = rho
    return out


def cumulant_to_central(c: list) -> list:
    """Inverse: ``κ̃_α = Σ_π Π_{B∈π} C̃_B`` (cumulant.py:747-815)."""
    rho = c[0]
    idx = {d: i for i, d in enumerate(DEGREES_27)}
    ct = [v / rho for v in c]
    out = list(c)
    for i, deg in enumerate(DEGREES_27):
        if sum(deg) < 2:
            continue
        acc = 0.0
        for part in _partitions_for(deg):
            prod = 1.0
            for b in part:
                prod = prod * ct[idx[b]]
            acc = acc + prod
        out[i] = acc * rho
    out[0] = rho
    return out


def _factorial(n: int) -> int:
    r = 1
    for i in range(2, n + 1):
        r *= i
    return r


def collide_cumulant_geier_d3q27_les(f, tau, omega_b=1.0, omega_odd=1.0, omega_even=1.0, C_s=0.0):
    """cumulant.py:1101-1110: with C_s > 0 the rate comes from the domain-mean Smagorinsky τ_eff
    clamped to [τ, 10 τ], as in the legacy operator."""
    if C_s > 0:
        tau = max(tau, min(smagorinsky_tau_mean(f, tau, C_s), 10.0 * tau))
    return collide_cumulant_geier_d3q27(f, tau, omega_b, omega_odd, omega_even)


def collide_cumulant_geier_d3q27(f, tau, omega_b=1.0, omega_odd=1.0, omega_even=1.0, *, omega=None):
    """Geier 2015 cumulant collision on the full f (cumulant.py:1050-1156).

    ``C* = C − s(C − C_eq)`` with ``C_eq`` from the equilibrium pushed through
    the same transform; trace/deviatoric split at second order.  The
    domain-averaged Smagorinsky path (C_s > 0) is not restated.
    """
    lat = D3Q27
    if f.shape[0] != 27:
        raise ValueError("D3Q27 populations required")
    if omega is None:
        omega = 1.0 / tau
    dt = f.dtype
    m_mat, m_inv = moment_matrix(lat, DEGREES_27)
    m_mat, m_inv = m_mat.astype(dt), m_inv.astype(dt)
    rho, ux, uy, uz = macroscopic(f)
    feq = equilibrium(lat, rho, ux, uy, uz)
    shape = f.shape
    mv = m_mat @ f.reshape(27, -1)
    mveq = m_mat @ feq.reshape(27, -1)
    m = [mv[i].reshape(shape[1:]) for i in range(27)]
    meq = [mveq[i].reshape(shape[1:]) for i in range(27)]
    c = central_to_cumulant(to_central(m, DEGREES_27, ux, uy, uz))
    ceq = central_to_cumulant(to_central(meq, DEGREES_27, ux, uy, uz))
    d = [a - b for a, b in zip(c, ceq)]
    cs = list(c)
    tr = d[4] + d[5] + d[6]
    for i in (4, 5, 6):
        cs[i] = c[i] - (omega * (d[i] - tr / 3.0) + omega_b * tr / 3.0)
    for i in (7, 8, 9):
        cs[i] = c[i] - omega * d[i]
    for i in range(10, 17):
        cs[i] = c[i] - omega_odd * d[i]
    for i in range(17, 27):
        cs[i] = c[i] - omega_even * d[i]
    ms = to_raw(cumulant_to_central(cs), DEGREES_27, ux, uy, uz)
    ms = np.stack([np.broadcast_to(v, shape[1:]) for v in ms]).reshape(27, -1)
    return (m_inv @ ms).reshape(shape)


# ---------------------------------------------------------------------------
# boundaries, forces, mass
# ---------------------------------------------------------------------------


def sphere_mask(nx, ny, nz, cx, cy, cz, radius) -> np.ndarray:
    """``(x−cx)² + (y−cy)² + (z−cz)² ≤ r²`` in float32 (boundaries3d.py:29-49): the centre rounded to
    float32, each square and sum rounded in turn, and r² formed in double and rounded to float32
    (torch compares a float32 tensor with a Python float in float32)."""
    zz, yy, xx = np.meshgrid(
        np.arange(nz, dtype=np.float32), np.arange(ny, dtype=np.float32),
        np.arange(nx, dtype=np.float32), indexing="ij",
    )
    d = (xx - np.float32(cx)) ** 2 + (yy - np.float32(cy)) ** 2 + (zz - np.float32(cz)) ** 2
    return d <= np.float32(float(radius) * float(radius))


def obstacle_force(f_streamed: np.ndarray, solid: np.ndarray):
    """Ladd momentum exchange ``F = 2 Σ_solid Σ_q c_q f`` (obstacles.py:91-130), float64 sum."""
    lat = _lat(f_streamed)
    fs = f_streamed.astype(np.float64) * solid[None]
    return tuple(2.0 * float((lat.c[:, a].reshape(-1, *([1] * (fs.ndim - 1))) * fs).sum()) for a in range(3))


def channel_bc(f: np.ndarray, u_in: float, solid: np.ndarray | None = None, wall: np.ndarray | None = None):
    """Equilibrium inlet at x=0 (ρ from x=1), copy outlet at x=nx−1, then bounce-back.

    boundaries3d.py:407-443 (D3Q19 / D3Q27 fields ``(Q, nz, ny, nx)``).
    """
    lat = _lat(f)
    f = f.copy()
    rho, ux, uy, uz = macroscopic(f)
    ux[:, :, 0] = u_in
    uy[:, :, 0] = 0.0
    uz[:, :, 0] = 0.0
    rho[:, :, 0] = rho[:, :, 1]
    f[:, :, :, 0] = equilibrium(lat, rho[:, :, 0], ux[:, :, 0], uy[:, :, 0], uz[:, :, 0])
    f[:, :, :, -1] = f[:, :, :, -2]
    if wall is not None:
        f = bounce_back(f, wall)
    if solid is not None:
        f = bounce_back(f, solid)
    return f


def zou_he_moving_lid(f: np.ndarray, u_lid: float) -> np.ndarray:
    """Zou-He lid on the y = ny−1 plane, ux = u_lid (boundaries3d.py:324-404)."""
    fl = f[:, :, -1, :]
    s0 = sum(fl[i] for i in (0, 1, 2, 5, 6, 11, 12, 13, 14))
    sp = sum(fl[i] for i in (3, 7, 10, 15, 17))
    rho = s0 + 2.0 * sp
    dx = rho * u_lid - ((fl[1] - fl[2] + fl[11] - fl[12] + fl[13] - fl[14]) + (fl[7] - fl[10]))
    sz = fl[5] - fl[6] + fl[11] - fl[12] - fl[13] + fl[14]
    out = f.copy()
    out[4, :, -1, :] = fl[3]
    out[8, :, -1, :] = 0.5 * (fl[7] + fl[10]) - 0.5 * dx
    out[9, :, -1, :] = 0.5 * (fl[7] + fl[10]) + 0.5 * dx
    out[16, :, -1, :] = fl[15] + 0.5 * sz
    out[18, :, -1, :] = fl[17] - 0.5 * sz
    return out


def correct_mass(f: np.ndarray, target_mass: float) -> np.ndarray:
    """Global rescale to ``target_mass`` (solver.py:420-437, solver3d.py:287-300)."""
    cur = float(f.astype(np.float64).sum())
    if abs(cur) < 1e-30:
        return f
    return (f * (target_mass / cur)).astype(f.dtype)


# ---------------------------------------------------------------------------
# fused obstacle step (kernel semantics)
# ---------------------------------------------------------------------------


def fused_obstacle_step(
    f: np.ndarray,
    tau: float,
    solid: np.ndarray | None = None,
    *,
    collision: str = "BGK",
    cs_delta_sq: float = 0.0,
    tau_eff: np.ndarray | None = None,
    with_force: bool = False,
):
    """Pull-stream + full-way BB + BGK/CM/CUMULANT with ω dispatch.

    triton_fused_obstacle.py:1035-1521: ``f_eff = where(solid, f[opp, x+c],
    f[q, x−c])``; ω = 1/τ_eff (external), the clamped Hou closure on the
    post-stream Π^neq (``cs_delta_sq > 0``), or 1/τ.  The force is the Ladd
    sum over solid cells of the post-stream, pre-bounce-back populations.
    Returns ``(f_post, force or None)``.
    """
    fin = stream(f)
    f_eff = fin if solid is None else bounce_back(fin, solid)
    feq, rho_safe, (ux, uy, uz) = _feq_safe(f_eff)
    fneq = f_eff - feq
    if tau_eff is not None:
        omega = 1.0 / tau_eff
    elif cs_delta_sq > 0.0:
        pxx, pyy, pzz, pxy, pxz, pyz = neq_stress(fneq)
        pn = np.sqrt(pxx * pxx + pyy * pyy + pzz * pzz + 2.0 * (pxy * pxy + pxz * pxz + pyz * pyz))
        omega = 1.0 / smagorinsky_tau(tau, pn, rho_safe, cs_delta_sq)
    else:
        omega = 1.0 / tau
    col = collision.upper()
    if col == "BGK":
        out = f_eff - omega * (f_eff - feq)
    elif col == "CM":
        out = collide_cascaded(f_eff, omega=omega)
    elif col == "CUMULANT":
        out = collide_cumulant_d3q19(f_eff, None, omega=omega)
    else:
        raise ValueError(f"unsupported collision {collision!r}")
    force = obstacle_force(fin, solid) if (with_force and solid is not None) else None
    return out.astype(f.dtype), force


__all__ = [
    "D2Q9", "D3Q19", "D3Q27", "equilibrium", "macroscopic", "neq_stress", "stream",
    "bounce_back", "fullway_pull", "collide_bgk", "fused_step", "smagorinsky_tau",
    "collide_smagorinsky_bgk", "collide_mrt", "mrt_matrices", "collide_cascaded",
    "collide_cumulant_d3q19", "collide_cumulant_geier_d3q27", "central_to_cumulant",
    "cumulant_to_central", "sphere_mask", "obstacle_force", "channel_bc",
    "zou_he_moving_lid", "correct_mass", "fused_obstacle_step",
]


# ---------------------------------------------------------------------------
# lid-driven cavity step (kernel semantics)
# ---------------------------------------------------------------------------


def cavity_walls(edge: int) -> np.ndarray:
    """Five stationary walls of a cubic cavity: x = 0, x = L−1, y = 0, z = 0, z = L−1
    (cases/cavity.py:72-79 walls, closed in z)."""
    m = np.zeros((edge, edge, edge), dtype=bool)
    m[:, :, 0] = m[:, :, -1] = True
    m[:, 0, :] = True
    m[0] = m[-1] = True
    return m


def cavity_step(f: np.ndarray, tau: float, u_lid: float) -> np.ndarray:
    """Pull stream, Zou-He lid on the fluid cells of y = L−1 (boundaries3d.py:324-404), full-way
    bounce-back on the walls, BGK with the fused kernels' floored-density equilibrium."""
    walls = cavity_walls(f.shape[1])
    fs = stream(f)
    fin = np.where(walls[None], fs, zou_he_moving_lid(fs, u_lid))
    fin = bounce_back(fin, walls)
    feq, _, _ = _feq_safe(fin)
    return fin - (1.0 / tau) * (fin - feq)


# ---------------------------------------------------------------------------
# D2Q9 central-moment ("cumulant") collision
# ---------------------------------------------------------------------------


def collide_cumulant_d2q9(f, tau, omega_b=1.0, omega_3=1.0, omega_4=1.0):
    """cumulant.py:116-311: raw moments m_pq (p, q ≤ 2) → central moments κ_pq by the binomial
    shift with (u_x, u_y) → relax (deviatoric stress at 1/τ, trace at ω_b, κ21/κ12 at ω_3, κ22 at
    ω_4) towards the central moments of the discrete equilibrium → shift back → populations."""
    lat = D2Q9
    rho, ux, uy, _ = macroscopic(f)
    feq = equilibrium(lat, rho, ux, uy)
    cx = lat.c[:, 0].astype(float).reshape(9, 1, 1)
    cy = lat.c[:, 1].astype(float).reshape(9, 1, 1)
    orders = [(2, 0), (0, 2), (1, 1), (2, 1), (1, 2), (2, 2)]

    def central(g):
        return {pq: (g * (cx - ux) ** pq[0] * (cy - uy) ** pq[1]).sum(0) for pq in orders}

    k, ke = central(f), central(feq)
    om = 1.0 / tau
    ks = {pq: k[pq] - om * (k[pq] - ke[pq]) for pq in ((2, 0), (0, 2), (1, 1))}
    t, te = k[(2, 0)] + k[(0, 2)], ke[(2, 0)] + ke[(0, 2)]
    d = 0.5 * ((t - omega_b * (t - te)) - (ks[(2, 0)] + ks[(0, 2)]))
    ks[(2, 0)] = ks[(2, 0)] + d
    ks[(0, 2)] = ks[(0, 2)] + d
    for pq, w in (((2, 1), omega_3), ((1, 2), omega_3), ((2, 2), omega_4)):
        ks[pq] = k[pq] - w * (k[pq] - ke[pq])
    # populations from the 9 central moments: solve the 9x9 system per cell
    # (rows: κ00 = ρ, κ10 = κ01 = 0 and the six relaxed moments)
    cols = [np.ones_like(cx * ux), cx - ux, cy - uy] + [(cx - ux) ** p * (cy - uy) ** q for p, q in orders]
    A = np.stack([np.broadcast_to(c, f.shape) for c in cols])  # (9 moments, 9 populations, ny, nx)
    b = np.stack([rho, np.zeros_like(rho), np.zeros_like(rho)] + [ks[pq] for pq in orders])
    A2 = np.moveaxis(A, (0, 1), (-2, -1))
    b2 = np.moveaxis(b, 0, -1)[..., None]
    return np.moveaxis(np.linalg.solve(A2, b2)[..., 0], -1, 0)


# ---------------------------------------------------------------------------
# legacy D3Q27 "cumulant" (collide_cumulant_d3q27)
# ---------------------------------------------------------------------------


def legacy_pseudo_cumulants(k: list) -> list:
    """cumulant.py:470-577: Geier's partition terms of orders 4-6 without the 1/ρ factors,
    applied to the central moments of f_neq; orders ≤ 3 pass through.  DEGREES_27 indices."""
    c = list(k)
    k200, k020, k002, k110, k101, k011 = k[4], k[5], k[6], k[7], k[8], k[9]
    c[17] = k[17] - k200 * k020 - 2.0 * k110 * k110
    c[18] = k[18] - k200 * k002 - 2.0 * k101 * k101
    c[19] = k[19] - k020 * k002 - 2.0 * k011 * k011
    c[20] = k[20] - k200 * k011 - 2.0 * k101 * k110
    c[21] = k[21] - k020 * k101 - 2.0 * k110 * k011
    c[22] = k[22] - k002 * k110 - 2.0 * k101 * k011
    c[23] = k[23] - k200 * k[13] - 2.0 * k110 * k[16]
    c[24] = k[24] - k200 * k[15] - 2.0 * k101 * k[16]
    c[25] = k[25] - k020 * k[14] - 2.0 * k011 * k[16]
    c[26] = (k[26] - k200 * c[19] - k020 * c[18] - k002 * c[17]
             - 2.0 * (k110 * c[22] + k101 * c[21] + k011 * c[20])
             + 2.0 * k200 * k020 * k002 + 4.0 * k110 * k101 * k011
             + 2.0 * k110 * k110 * k002 + 2.0 * k101 * k101 * k020 + 2.0 * k011 * k011 * k200)
    return c


def legacy_central_moments(c: list) -> list:
    """cumulant.py:580-651: the inverse map of ``legacy_pseudo_cumulants``."""
    k = list(c)
    a, b, d, xy, xz, yz = c[4], c[5], c[6], c[7], c[8], c[9]
    k[17] = c[17] + a * b + 2.0 * xy * xy
    k[18] = c[18] + a * d + 2.0 * xz * xz
    k[19] = c[19] + b * d + 2.0 * yz * yz
    k[20] = c[20] + a * yz + 2.0 * xz * xy
    k[21] = c[21] + b * xz + 2.0 * xy * yz
    k[22] = c[22] + d * xy + 2.0 * xz * yz
    k[23] = c[23] + a * c[13] + 2.0 * xy * c[16]
    k[24] = c[24] + a * c[15] + 2.0 * xz * c[16]
    k[25] = c[25] + b * c[14] + 2.0 * yz * c[16]
    k[26] = (c[26] + a * c[19] + b * c[18] + d * c[17] + 2.0 * (xy * c[22] + xz * c[21] + yz * c[20])
             - 2.0 * a * b * d - 4.0 * xy * xz * yz - 2.0 * xy * xy * d - 2.0 * xz * xz * b - 2.0 * yz * yz * a)
    return k


def smagorinsky_tau_mean(f: np.ndarray, tau: float, C_s: float) -> float:
    """Domain mean of the clamped per-cell Smagorinsky τ_eff (cumulant.py:986-997 before its clamp)."""
    lat = _lat(f)
    rho, ux, uy, uz = macroscopic(f)
    fneq = f - equilibrium(lat, rho, ux, uy, uz)
    return float(smagorinsky_tau(tau, pi_norm(fneq), rho, C_s * C_s).mean())


def collide_cumulant_d3q27(f, tau, omega_b=1.0, omega_odd=1.0, omega_even=1.0, C_s=0.0):
    """cumulant.py:935-1047: f_neq raw moments → central → pseudo-cumulants → relax → back.
    C_s > 0: ω from the domain-mean Smagorinsky τ_eff, clamped to [τ, 10 τ]."""
    lat = _lat(f)
    omega = 1.0 / tau
    if C_s > 0:
        te = smagorinsky_tau_mean(f, tau, C_s)
        omega = 1.0 / max(tau, min(te, 10.0 * tau))
    rho, ux, uy, uz = macroscopic(f)
    feq = equilibrium(lat, rho, ux, uy, uz)
    M, Mi = moment_matrix(lat, DEGREES_27)
    sh = f.shape[1:]
    m = list((M @ (f - feq).reshape(27, -1)).reshape((27,) + sh))
    k = legacy_pseudo_cumulants(to_central(m, DEGREES_27, ux, uy, uz))
    _relax_second(k, omega, omega_b)
    for i in range(10, 17):
        k[i] = (1.0 - omega_odd) * k[i]
    for i in range(17, 27):
        k[i] = (1.0 - omega_even) * k[i]
    m = to_raw(legacy_central_moments(k), DEGREES_27, ux, uy, uz)
    return feq + (Mi @ np.stack(m).reshape(27, -1)).reshape((27,) + sh)


def far_field_6face(f: np.ndarray, u_in: float, uy: float = 0.0, uz: float = 0.0) -> np.ndarray:
    """triton_fused_obstacle.py:2386-2471 (apply_far_field_bc_6face): equilibrium (ρ = 1) on x = 0, copy
    x = nx−2 onto x = nx−1, equilibrium on y = 0 / ny−1 and, only when nz > 4, on z = 0 / nz−1; in that
    order, out of place here."""
    lat = _lat(f)
    one = np.ones((1, 1, 1))
    feq = equilibrium(lat, one, one * u_in, one * uy, one * uz)[:, 0, 0, 0][:, None, None]
    f = f.copy()
    nz, ny, nx = f.shape[1:]
    f[:, :, :, 0] = feq
    if nx >= 2:
        f[:, :, :, -1] = f[:, :, :, -2]
    f[:, :, 0, :] = feq
    if ny >= 2:
        f[:, :, -1, :] = feq
    if nz > 4:
        f[:, 0, :, :] = feq
        f[:, -1, :, :] = feq
    return f


def far_field_bc(f: np.ndarray, u_in: float, obstacle: np.ndarray | None = None, uy: float = 0.0, uz: float = 0.0,
                 bc_config: dict | None = None) -> np.ndarray:
    """boundaries3d.py:446-537: free-stream equilibrium (ρ = 1) on the inlet x = 0 and the selected
    lateral faces, copy outlet x = nx−1, then bounce-back in the obstacle; face order as there."""
    lat = _lat(f)
    one = np.ones((1, 1, 1))
    feq = equilibrium(lat, one, one * u_in, one * uy, one * uz)[:, 0, 0, 0]
    f = f.copy()
    if bc_config is None:
        ff, periodic = {"y-", "y+", "z-", "z+"}, set()
    else:
        ff, periodic = set(bc_config.get("far_field_faces", [])), set(bc_config.get("periodic_faces", []))
    if "x-" not in periodic:
        f[:, :, :, 0] = feq[:, None, None]
    if "x+" not in periodic:
        f[:, :, :, -1] = f[:, :, :, -2]
    if "y-" in ff and "y-" not in periodic:
        f[:, :, 0, :] = feq[:, None, None]
    if "y+" in ff and "y+" not in periodic:
        f[:, :, -1, :] = feq[:, None, None]
    if "z-" in ff and "z-" not in periodic:
        f[:, 0, :, :] = feq[:, None, None]
    if "z+" in ff and "z+" not in periodic:
        f[:, -1, :, :] = feq[:, None, None]
    if obstacle is not None:
        f = bounce_back(f, obstacle)
    return f
