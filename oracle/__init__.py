"""CPU oracle for the per-step lattice update — TEST INFRASTRUCTURE ONLY.

This package restates the reference TensorLBM algorithms for the hot path
(equilibrium, macroscopic moments, BGK/MRT/cascaded/cumulant collisions,
periodic pull streaming, full-way bounce-back, Smagorinsky closure, Ladd
momentum-exchange force, channel / moving-lid boundaries) in plain numpy,
plus a small OpenMP C restatement of the fused step (``oracle/c``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker.  The product package
``tensorlbm_amd`` never imports, calls or links anything in here; its GPU
path fails loudly when the HIP library is missing.

Each function cites the reference file:line it follows.  How each function
is pinned (golden vectors, closed forms, invariants, brute force) is listed
in ``tests/test_oracle_*.py``.  Functions with no independent pin are
marked "parity unpinned" in their docstring.
"""

from . import lattice, ref  # noqa: F401
