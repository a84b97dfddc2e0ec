"""Generated moment tables (host logic; no GPU).

The kernels reach the MRT basis through the monomial moments and two sparse
recombinations, M = T CM and M^-1 = CM^-1 TI.  In exact rationals: those
identities hold for D3Q19 and D3Q27, T and TI are the sparse matrices the
kernels count on, and the committed tables.inc is what the generator writes.
"""

from __future__ import annotations

import importlib.util
from fractions import Fraction
from pathlib import Path

import pytest

CSRC = Path(__file__).resolve().parents[1] / "tensorlbm_amd" / "csrc"


def _gen():
    spec = importlib.util.spec_from_file_location("gen_tables", CSRC / "gen_tables.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


G = _gen()


def test_committed_tables_are_the_generated_ones():
    assert (CSRC / "tables.inc").read_text() == G.render()


@pytest.mark.parametrize("q", [19, 27])
def test_mrt_factors_through_the_monomial_moments(q):
    mrt, c, deg = {19: (G.mrt_d3q19(), G.D3Q19_C, G.DEG19), 27: (G.mrt_d3q27(), G.D3Q27_C, G.DEG27)}[q]
    cm = G.monomial(c, deg)
    t = G.matmul(mrt, G.inverse(cm))
    ti = G.matmul(cm, G.inverse(mrt))
    assert G.matmul(t, cm) == [[Fraction(v) for v in row] for row in mrt]
    eye = [[Fraction(int(i == j)) for j in range(q)] for i in range(q)]
    assert G.matmul(G.matmul(G.inverse(cm), ti), mrt) == eye  # CM^-1 TI = M^-1
    nnz = lambda m: sum(v != 0 for row in m for v in row)  # noqa: E731
    assert nnz(t) + nnz(ti) < (nnz(mrt) + nnz(G.inverse(mrt))) // 3


def test_d3q19_moments_are_d3q27_moments_without_corners():
    """Every D3Q19 exponent is a D3Q27 one, and the D3Q27 moments a corner-free distribution lacks in
    D3Q19 (all three exponents nonzero) vanish for it: the D3Q27 sweeps serve D3Q19 exactly."""
    assert set(G.DEG19) <= set(G.DEG27)
    for deg in set(G.DEG27) - set(G.DEG19):
        assert all(e > 0 for e in deg)
        row = G.monomial(G.D3Q19_C, [deg])[0]
        assert all(v == 0 for v in row)


def _equilibrium_cumulants_symbolic():
    """Cumulants per unit density of the second-order D3Q27 equilibrium, derived symbolically the way the
    kernels used to compute them: raw moments of w_q (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2), binomial shift
    to central moments, the set-partition cumulant map of the generated tables."""
    sp = pytest.importorskip("sympy")
    from itertools import product
    from math import comb

    from oracle import ref

    ux, uy, uz = sp.symbols("ux uy uz")
    u = (ux, uy, uz)
    c = ref.D3Q27.c
    w = [sp.Rational(8, 27) if not any(v) else sp.Rational(2, 27) if sum(map(abs, v)) == 1 else
         sp.Rational(1, 54) if sum(map(abs, v)) == 2 else sp.Rational(1, 216) for v in c.tolist()]
    usq = ux**2 + uy**2 + uz**2
    feq = [w[i] * (1 + 3 * sum(int(c[i, d]) * u[d] for d in range(3))
                   + sp.Rational(9, 2) * sum(int(c[i, d]) * u[d] for d in range(3)) ** 2
                   - sp.Rational(3, 2) * usq) for i in range(27)]
    raw = {e: sp.expand(sum(feq[i] * int(c[i, 0]) ** e[0] * int(c[i, 1]) ** e[1] * int(c[i, 2]) ** e[2]
                            for i in range(27))) for e in G.DEG27}
    cen = {}
    for e in G.DEG27:
        acc = 0
        for k in product(*[range(ed + 1) for ed in e]):
            coef = 1
            for d in range(3):
                coef *= comb(e[d], k[d]) * (-u[d]) ** (e[d] - k[d])
            acc += coef * raw[tuple(k)]
        cen[e] = sp.expand(acc)
    kt = [cen[e] for e in G.DEG27]
    out = []
    for i, dg in enumerate(G.DEG27):
        if sum(dg) < 2:
            out.append(kt[i])
            continue
        acc = 0
        for key, coef in G.cumulant_terms(dg, False).items():
            term = sp.Rational(coef.numerator, coef.denominator)
            for k in key:
                term *= kt[k]
            acc += term
        out.append(sp.expand(acc))
    return sp, u, raw, out


def test_geier_equilibrium_closed_forms():
    """The closed forms the Geier collision uses (collide.hpp: f27::eq_moments, eq_cumulants27)."""
    sp, (x, y, z), raw, cum = _equilibrium_cumulants_symbolic()
    m1d = lambda n: 1 if n == 0 else (0 if n % 2 else sp.Rational(1, 3))  # noqa: E731
    s3 = lambda e: m1d(e[0]) * m1d(e[1]) * m1d(e[2])  # noqa: E731
    u = (x, y, z)
    for e in G.DEG27:  # raw moments per unit density
        closed = s3(e) * (1 - sp.Rational(3, 2) * (x**2 + y**2 + z**2))
        for d in range(3):
            closed += 3 * s3(tuple(e[k] + (k == d) for k in range(3))) * u[d]
            for f in range(3):
                closed += sp.Rational(9, 2) * s3(tuple(e[k] + (k == d) + (k == f) for k in range(3))) * u[d] * u[f]
        assert sp.expand(closed - raw[e]) == 0, e
    want = [1, 0, 0, 0, sp.Rational(1, 3), sp.Rational(1, 3), sp.Rational(1, 3), 0, 0, 0,
            -x * x * y, -x * x * z, -y * y * x, -y * y * z, -z * z * x, -z * z * y, -x * y * z,
            3 * x * x * y * y, 3 * x * x * z * z, 3 * y * y * z * z, 3 * x * y * z * x, 3 * x * y * z * y,
            3 * x * y * z * z, -6 * x * y * z * x * y, -6 * x * y * z * x * z, -6 * x * y * z * y * z, 0]
    for i, (a, b) in enumerate(zip(cum, want)):
        assert sp.expand(a - b) == 0, (i, G.DEG27[i], a)
