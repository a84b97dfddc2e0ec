"""``torch.ops.tlbm``: the C ABI registered with torch's dispatcher (``libtlbm_torch.so``).

The ctypes layer (``_native``, ``ops``, ``fused``) is the binding a reference maintainer adds;
these operators are the same entry points as dispatcher ops, for callers that need torch to see
them: ``torch.compile`` (Meta kernels give the output shapes), graph capture and the profiler.

* ``tlbm.equilibrium(Q, rho, ux, uy, uz=None)`` → ``tlbm_equilibrium``
* ``tlbm.macroscopic(f)`` → ``[rho, ux, uy(, uz)]`` (``tlbm_macroscopic``)
* ``tlbm.stream(f)`` → ``tlbm_stream``
* ``tlbm.collide(f, kind, tau, rates=[], tau_mode=0, cs_delta_sq=0., tau_eff=None)`` → ``tlbm_collide``
* ``tlbm.step(f, kind, tau, rates=[], tau_mode=0, cs_delta_sq=0., tau_eff=None, solid=None,
  bc_kind=0, bc_u=[0, 0, 0], bc_faces=0, force=None, deviation=False, steps=1)`` → ``tlbm_step`` /
  ``tlbm_step2``; ``force`` (3 float64 entries) is added to.

``kind``, ``tau_mode`` and ``bc_kind`` take the ``TLBM_*`` codes of ``include/tlbm_hip.h``
(``_native.COLL_*``, ``TAU_*``, ``BC_*``).  Outputs are new contiguous tensors on the input's
device, computed on torch's current stream.  Errors raise the reference's exception types.
"""

from __future__ import annotations

from pathlib import Path

import torch

LIB_PATH = Path(__file__).with_name("libtlbm_torch.so")
_loaded = False


def load():
    """Register the operators (once per process) and return the ``torch.ops.tlbm`` namespace."""
    global _loaded
    if not _loaded:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH.name} is not built; run `python -m tensorlbm_amd.csrc.build`")
        torch.ops.load_library(str(LIB_PATH))
        _loaded = True
    return torch.ops.tlbm


__all__ = ["load", "LIB_PATH"]
