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
