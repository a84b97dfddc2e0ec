"""Velocity sets of the oracle (test infrastructure only).

Orderings follow the reference exactly:

* D2Q9  — ``src/tensorlbm/d2q9.py:7-27`` (rest, +x, +y, -x, -y, diagonals).
* D3Q19 — ``src/tensorlbm/d3q19.py:6-67`` (rest, ±x, ±y, ±z, then xy, xz,
  yz edge pairs; note OPPOSITE[q] != 18 - q).
* D3Q27 — ``src/tensorlbm/d3q27.py:32-84`` (this is NOT the z,y,x loop order
  of ``core/lattice.py``).

Weights are exact rationals evaluated in binary64 (the reference keeps
``W_EXACT64`` for D3Q19/D3Q27; its D2Q9 table is float32-only,
``d2q9.py:22-25``, which only matters for fp64 runs).

D2Q9 is carried as a 3-D velocity set with ``cz = 0`` so that a 2-D field
``(9, ny, nx)`` is the ``nz = 1`` case of the 3-D code.
"""

from __future__ import annotations

from dataclasses import dataclass
from fractions import Fraction

import numpy as np


@dataclass(frozen=True)
class Lattice:
    name: str
    c: np.ndarray  # (Q, 3) int64
    w_frac: tuple  # exact Fractions
    dim: int

    @property
    def q(self) -> int:
        return self.c.shape[0]

    @property
    def w(self) -> np.ndarray:
        return np.array([float(v) for v in self.w_frac], dtype=np.float64)

    @property
    def opposite(self) -> np.ndarray:
        rows = [tuple(r) for r in self.c.tolist()]
        return np.array([rows.index(tuple(-v for v in r)) for r in rows], dtype=np.int64)


_D2Q9_C = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]
D2Q9 = Lattice(
    "D2Q9",
    np.array([(cx, cy, 0) for cx, cy in _D2Q9_C], dtype=np.int64),
    tuple([Fraction(4, 9)] + [Fraction(1, 9)] * 4 + [Fraction(1, 36)] * 4),
    2,
)

D3Q19 = Lattice(
    "D3Q19",
    np.array(
        [
            (0, 0, 0),
            (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
            (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0),
            (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1),
            (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1),
        ],
        dtype=np.int64,
    ),
    tuple([Fraction(1, 3)] + [Fraction(1, 18)] * 6 + [Fraction(1, 36)] * 12),
    3,
)

D3Q27 = Lattice(
    "D3Q27",
    np.array(
        [
            (0, 0, 0),
            (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
            (1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0),
            (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
            (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1),
            (1, 1, 1), (-1, 1, 1), (1, -1, 1), (-1, -1, 1),
            (1, 1, -1), (-1, 1, -1), (1, -1, -1), (-1, -1, -1),
        ],
        dtype=np.int64,
    ),
    tuple(
        [Fraction(8, 27)] + [Fraction(2, 27)] * 6 + [Fraction(1, 54)] * 12 + [Fraction(1, 216)] * 8
    ),
    3,
)

BY_NAME = {"D2Q9": D2Q9, "D3Q19": D3Q19, "D3Q27": D3Q27}


def for_q(q: int) -> Lattice:
    for lat in (D2Q9, D3Q19, D3Q27):
        if lat.q == q:
            return lat
    raise ValueError(f"no lattice with Q={q}")
