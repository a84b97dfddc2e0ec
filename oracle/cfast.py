"""ctypes wrapper of the C oracle ``oracle/c/lbm_oracle.c`` (test infrastructure).

``fused_bgk(f, tau, solid=None)`` = ``collide_bgk(stream(f))`` with the fused
kernel's moment convention, OpenMP-parallel on the host cores.
"""

from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

from .lattice import for_q

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_build" / "liboracle_c.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_HERE / "c")], check=True)
    return LIB_PATH


def available() -> bool:
    return LIB_PATH.is_file()


def _load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(str(LIB_PATH))
        for name in ("oracle_fused_bgk_f32", "oracle_fused_bgk_f64"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_int] * 3 + [ctypes.c_double]
        lib.oracle_num_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def num_threads() -> int:
    return int(_load().oracle_num_threads())


def fused_bgk(f: np.ndarray, tau: float, solid: np.ndarray | None = None, out: np.ndarray | None = None) -> np.ndarray:
    lat = for_q(f.shape[0])
    f4 = f[:, None] if f.ndim == 3 else f
    f4 = np.ascontiguousarray(f4)
    if out is None:
        out = np.empty_like(f4)
    c = np.ascontiguousarray(lat.c.astype(np.int32))
    opp = np.ascontiguousarray(lat.opposite.astype(np.int32))
    w = np.ascontiguousarray(lat.w)
    sp = None
    if solid is not None:
        sp = np.ascontiguousarray(solid.reshape(f4.shape[1:]).astype(np.int8))
    fn = _load().oracle_fused_bgk_f32 if f4.dtype == np.float32 else _load().oracle_fused_bgk_f64
    rc = fn(lat.q, c.ctypes.data, opp.ctypes.data, w.ctypes.data, f4.ctypes.data, out.ctypes.data,
            None if sp is None else sp.ctypes.data, *f4.shape[1:], float(tau))
    if rc != 0:
        raise RuntimeError(f"oracle_fused_bgk failed ({rc})")
    return out[:, 0] if f.ndim == 3 else out
