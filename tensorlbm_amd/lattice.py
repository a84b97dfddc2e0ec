"""Lattice constant tables with the reference's module-level names.

``D2Q9.C/W/OPPOSITE`` mirror ``src/tensorlbm/d2q9.py:7-27``,
``D3Q19.*`` mirror ``src/tensorlbm/d3q19.py:6-67`` (with ``W_EXACT64``) and
``D3Q27.*`` mirror ``src/tensorlbm/d3q27.py:32-84``.  The HIP kernels carry
the same tables as compile-time constants (``csrc/lattice.hpp``); the test
suite checks the two agree.
"""

from __future__ import annotations

from types import SimpleNamespace

import torch

_D2Q9_C = [[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1], [1, 1], [-1, 1], [-1, -1], [1, -1]]
_D2Q9_W = [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4

_D3Q19_C = [
    [0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1],
    [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1],
]
_D3Q19_W = [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12

_D3Q27_C = [
    [0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0], [1, 0, 1], [-1, 0, 1], [1, 0, -1],
    [-1, 0, -1], [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1], [1, 1, 1], [-1, 1, 1],
    [1, -1, 1], [-1, -1, 1], [1, 1, -1], [-1, 1, -1], [1, -1, -1], [-1, -1, -1],
]
_D3Q27_W = [8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8


def _opposite(rows):
    return [rows.index([-v for v in r]) for r in rows]


def _table(rows, weights):
    return SimpleNamespace(
        C=torch.tensor(rows, dtype=torch.int64),
        W=torch.tensor(weights, dtype=torch.float32),
        W_EXACT64=torch.tensor(weights, dtype=torch.float64),
        OPPOSITE=torch.tensor(_opposite(rows), dtype=torch.int64),
        Q=len(rows),
    )


D2Q9 = _table(_D2Q9_C, _D2Q9_W)
D3Q19 = _table(_D3Q19_C, _D3Q19_W)
D3Q27 = _table(_D3Q27_C, _D3Q27_W)

BY_Q = {9: D2Q9, 19: D3Q19, 27: D3Q27}
