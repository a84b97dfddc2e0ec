"""Write input/output vectors of reference functions that no closed form pins.

Run in a container that holds the reference checkout (read-only), never on a
GPU machine and never from a test:

    python tests/golden/make_reference_vectors.py /path/to/reference

It loads the reference's hot-path leaf modules (``src/tensorlbm/*.py``) under
a stub ``tensorlbm`` package whose ``__path__`` points at that directory, so
the package ``__init__`` (which needs Python >= 3.11) never runs.  Only these
functions are evaluated, on small seeded fp64 / fp32 states:

* ``cumulant.collide_cumulant_d3q27`` (cumulant.py:935-1047) -- the legacy
  pseudo-cumulant operator; its 5th/6th-order map and the domain-mean
  Smagorinsky rate are fixed only by its source.
* ``d3q27.collide_mrt27`` (d3q27.py:643-726) -- the Gram-Schmidt basis keeps
  rows that become zero by substituting standard basis vectors, so which
  directions relax at s_q and which at s_pi is fixed only by the source.
* ``cumulant.collide_cumulant_d3q19`` with the WALE (``C_w``) and Vreman
  (``C_v``) closures, with and without a solid mask and a wall velocity
  (cumulant.py:1159-1377; gradient stencils of turbulence.py).

Output: ``reference_vectors.npz`` next to this script.  Every array is data:
inputs, parameters and the reference's outputs.
"""

from __future__ import annotations

import importlib
import sys
import types
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "reference_vectors.npz"


def load_reference(root: Path):
    """Import the leaf modules without executing the package ``__init__``."""
    src = root / "src" / "tensorlbm"
    if not (src / "cumulant.py").is_file():
        raise SystemExit(f"no reference sources under {src}")
    pkg = types.ModuleType("tensorlbm")
    pkg.__path__ = [str(src)]
    sys.modules["tensorlbm"] = pkg
    mods = {}
    for name in ("d2q9", "d3q19", "d3q27", "cumulant"):
        mods[name] = importlib.import_module(f"tensorlbm.{name}")
    return mods


def seeded_state(seed: int, q: int, shape=(3, 4, 5), amp=0.08):
    """A random (rho, u) and a seeded relative non-equilibrium part for ``f = f_eq (1 + neq)``."""
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.05 * rng.uniform(-1, 1, shape)
    u = 0.06 * rng.uniform(-1, 1, (3,) + shape)
    return rho, u, rng.uniform(-1, 1, (q,) + shape) * amp


def populations(eq, seed: int, q: int, shape=(3, 4, 5), amp=0.08):
    """Strictly positive fp64 populations: the equilibrium at a random state times (1 + neq)."""
    import torch

    rho, u, neq = seeded_state(seed, q, shape, amp)
    feq = eq(*(torch.from_numpy(np.ascontiguousarray(a)) for a in (rho, u[0], u[1], u[2])))
    return feq + torch.from_numpy(neq) * feq


def main():
    import torch

    root = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    ref = load_reference(root)
    eq27 = ref["d3q27"].equilibrium27
    out: dict[str, np.ndarray] = {}

    # -- legacy D3Q27 pseudo-cumulant operator ------------------------------------------------------
    legacy_cases = [
        # tau, omega_b, omega_odd, omega_even, C_s
        (0.6, 1.0, 1.0, 1.0, 0.0),
        (0.55, 1.3, 1.1, 0.9, 0.0),
        (0.8, 0.7, 1.6, 1.25, 0.0),
        (0.52, 1.0, 1.0, 1.0, 0.17),
        (0.7, 1.2, 0.8, 1.4, 0.1),
    ]
    for i, case in enumerate(legacy_cases):
        f = populations(eq27, 100 + i, 27)
        tau, ob, oo, oe, cs = case
        res = ref["cumulant"].collide_cumulant_d3q27(f.clone(), tau, omega_b=ob, omega_odd=oo, omega_even=oe, C_s=cs)
        out[f"legacy27_{i}_in"] = f.numpy()
        out[f"legacy27_{i}_params"] = np.array(case, dtype=np.float64)
        out[f"legacy27_{i}_out"] = res.numpy()

    # -- D3Q27 MRT: every rate distinct (the reference caches the matrices in fp32) ----------------
    mrt_cases = [
        # tau, s_e, s_eps, s_q, s_pi
        (0.6, 1.19, 1.4, 1.2, None),
        (0.9, 0.3, 1.9, 0.7, 1.5),
        (0.55, 1.7, 0.45, 1.05, 0.8),
    ]
    for i, case in enumerate(mrt_cases):
        f = populations(eq27, 200 + i, 27, amp=0.02).to(torch.float32)
        tau, s_e, s_eps, s_q, s_pi = case
        res = ref["d3q27"].collide_mrt27(f.clone(), tau, s_e=s_e, s_eps=s_eps, s_q=s_q, s_pi=s_pi)
        out[f"mrt27_{i}_in"] = f.numpy()
        out[f"mrt27_{i}_params"] = np.array([tau, s_e, s_eps, s_q, np.nan if s_pi is None else s_pi])
        out[f"mrt27_{i}_out"] = res.numpy()

    # -- D3Q19 cumulant with the gradient closures (WALE, Vreman) ---------------------------------
    eq19 = ref["d3q19"].equilibrium3d
    sgs_cases = [
        # shape, tau, omega_b, omega_even, C_w, C_v, solid fraction, wall velocity
        ((4, 5, 6), 0.55, 1.0, 1.0, 0.5, 0.0, 0.0, (0.0, 0.0, 0.0)),
        ((4, 5, 6), 0.62, 1.2, 0.9, 0.0, 0.025, 0.0, (0.0, 0.0, 0.0)),
        ((5, 4, 7), 0.51, 1.0, 1.1, 0.5, 0.0, 0.25, (0.01, -0.02, 0.005)),
        ((5, 4, 7), 0.58, 0.8, 1.0, 0.0, 0.07, 0.25, (0.0, 0.03, 0.0)),
        ((2, 1, 5), 0.6, 1.0, 1.0, 0.5, 0.0, 0.0, (0.0, 0.0, 0.0)),
        ((3, 2, 1), 0.6, 1.0, 1.0, 0.0, 0.025, 0.0, (0.0, 0.0, 0.0)),
    ]
    for i, (shape, tau, ob, oe, cw, cv, frac, wall) in enumerate(sgs_cases):
        f = populations(eq19, 300 + i, 19, shape)
        mask = None
        if frac > 0.0:
            mask = torch.from_numpy(np.random.default_rng(400 + i).uniform(0, 1, shape) < frac)
        res = ref["cumulant"].collide_cumulant_d3q19(f.clone(), tau, omega_b=ob, omega_even=oe, C_w=cw, C_v=cv,
                                                      solid_mask=mask, wall_velocity=wall)
        out[f"sgs19_{i}_in"] = f.numpy()
        out[f"sgs19_{i}_params"] = np.array([tau, ob, oe, cw, cv, *wall], dtype=np.float64)
        out[f"sgs19_{i}_mask"] = (mask.numpy() if mask is not None else np.zeros(shape, bool))
        out[f"sgs19_{i}_has_mask"] = np.array(mask is not None)
        out[f"sgs19_{i}_out"] = res.numpy()

    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({len(out)} arrays)")


if __name__ == "__main__":
    main()
