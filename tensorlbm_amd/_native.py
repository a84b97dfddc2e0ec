"""ctypes binding of ``libtlbm_hip.so`` (declared in ``include/tlbm_hip.h``).

The library is loaded lazily on first use, after ``torch`` so that it binds to
the HIP runtime torch already holds.  There is no fallback: if the shared
object is missing or a call fails, an exception names the problem.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch

LIB_NAME = "libtlbm_hip.so"
LIB_PATH = Path(__file__).with_name(LIB_NAME)

# status codes (tlbm_hip.h)
OK = 0
ERR_INVALID = -1
ERR_SHAPE = -2
ERR_DTYPE = -3
ERR_UNSUPPORTED = -4
ERR_ALIAS = -5
ERR_HIP = -100

# lattices / storage / collision / tau modes / boundaries
D2Q9, D3Q19, D3Q27 = 9, 19, 27
STORE_F32, STORE_F64, STORE_F16, STORE_BF16, STORE_F16_DEV, STORE_BF16_DEV = range(6)
COLL_NONE, COLL_BGK, COLL_MRT, COLL_CM, COLL_CUMULANT, COLL_GEIER = range(6)
TAU_CONST, TAU_SMAG, TAU_FIELD, TAU_SMAG_RAW, TAU_MLP, TAU_FIELD64, TAU_MLP_BF16 = range(7)
SGS_WALE, SGS_VREMAN = 1, 2
BC_NONE, BC_CHANNEL_X, BC_LID_Y, BC_FAR_FIELD = range(4)


class Collision(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("tau_mode", ctypes.c_int32),
        ("eq_rho_floor", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("tau", ctypes.c_double),
        ("cs_delta_sq", ctypes.c_double),
        ("rate", ctypes.c_double * 8),
        ("tau_eff", ctypes.c_void_p),
        ("mlp", ctypes.c_void_p),
    ]


class Cavity(ctypes.Structure):
    _fields_ = [
        ("edge", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("tau", ctypes.c_double),
        ("offset", ctypes.c_int64),
        ("tiles_before", ctypes.c_int64),
    ]


CAVITIES_MAX = 128
MLP_NOALLOC_MAX_SAMPLES = 32768  # TLBM_MLP_NOALLOC_MAX_SAMPLES


class Boundary(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("faces", ctypes.c_int32),
        ("u", ctypes.c_double * 3),
    ]


FACE_YM, FACE_YP, FACE_ZM, FACE_ZP, FACE_XM_PERIODIC, FACE_XP_PERIODIC = 1, 2, 4, 8, 16, 32
_FACE_BITS = {"y-": FACE_YM, "y+": FACE_YP, "z-": FACE_ZM, "z+": FACE_ZP}


def far_field_faces(bc_config: dict | None) -> int:
    """TLBM_FACE_* bits of far_field_bc_3d's ``bc_config`` (boundaries3d.py:446-537; None: all four lateral faces)."""
    if bc_config is None:
        return FACE_YM | FACE_YP | FACE_ZM | FACE_ZP
    ff = set(bc_config.get("far_field_faces", []))
    periodic = set(bc_config.get("periodic_faces", []))
    bits = 0
    for face, bit in _FACE_BITS.items():
        if face in ff and face not in periodic:
            bits |= bit
    if "x-" in periodic:
        bits |= FACE_XM_PERIODIC
    if "x+" in periodic:
        bits |= FACE_XP_PERIODIC
    return bits


_VP = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64

_SIGNATURES = {
    "tlbm_version": ([], _I),
    "tlbm_last_error": ([], ctypes.c_char_p),
    "tlbm_status_string": ([_I], ctypes.c_char_p),
    "tlbm_step": ([_I, _I, _VP, _VP, _VP, _VP, _I, _I, _I, ctypes.POINTER(Collision), _VP,
                   ctypes.POINTER(Boundary), _VP, _VP], _I),
    "tlbm_step2": ([_I, _I, _VP, _VP, _VP, _VP, _I, _I, _I, ctypes.POINTER(Collision), _VP,
                    ctypes.POINTER(Boundary), _VP, _VP], _I),
    "tlbm_step2_supported": ([_I, _I, ctypes.POINTER(Collision), ctypes.POINTER(Boundary), _I], _I),
    "tlbm_stream": ([_I, _I, _VP, _VP, _VP, _VP, _I, _I, _I, _VP], _I),
    "tlbm_collide": ([_I, _I, _VP, _VP, _VP, _VP, _I, _I, _I, ctypes.POINTER(Collision), _VP], _I),
    "tlbm_equilibrium": ([_I, _I, _VP, _VP, _VP, _VP, _VP, _I64, _VP], _I),
    "tlbm_macroscopic": ([_I, _I, _VP, _VP, _VP, _VP, _VP, _I64, _VP], _I),
    "tlbm_bounce_back": ([_I, _I, _VP, _VP, _VP, _I64, _VP], _I),
    "tlbm_moment_sums": ([_I, _I, _VP, _VP, _I, _I, _I, _VP, _VP], _I),
    "tlbm_obstacle_force": ([_I, _I, _VP, _VP, _I, _I, _I, _VP, _VP, _VP], _I),
    "tlbm_scale": ([_I, _I, _VP, _VP, _I, _I, _I, _VP, _VP], _I),
    "tlbm_boundary_eager": ([_I, _VP, _VP, _I, _I, _I, ctypes.POINTER(Boundary), _VP, _VP, _VP], _I),
    "tlbm_sphere_mask": ([_I, _I, _I, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _VP,
                          _VP], _I),
    "tlbm_gradient_closure_tau": ([_I, _VP, _VP, _VP, _VP, _VP, _I, _I, _I, _I, ctypes.c_double, ctypes.c_double,
                                   _VP, _VP], _I),
    "tlbm_smagorinsky_tau_sum": ([_I, _I, _VP, _VP, _I, _I, _I, ctypes.c_double, ctypes.c_double, _VP, _VP], _I),
    "tlbm_mlp_sizes": ([_VP, _VP, _VP, _VP], _I),
    "tlbm_mlp_pack": ([_VP, _VP, _VP], _I),
    "tlbm_mlp_pack_bf16": ([_VP, _VP, _VP], _I),
    "tlbm_mlp_predict": ([_VP, _VP, _I64, _VP, _VP], _I),
    "tlbm_mlp_grad": ([_VP, _VP, _VP, _I64, _VP, _VP], _I),
    "tlbm_mlp_adam": ([_VP, _VP, _I, _VP, _VP, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                       ctypes.c_float, ctypes.c_float, _VP, _VP, _VP], _I),
    "tlbm_mlp_train_step_noalloc": ([_VP, _VP, _VP, _I64, _VP, _VP, _I, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float, _VP, _VP, _VP], _I),
    "tlbm_mlp2d_nparam": ([_VP], _I64),
    "tlbm_mlp2d_forward": ([_VP, _VP, _VP, _I64, _VP, _VP], _I),
    "tlbm_strain2d": ([_VP, _VP, _I, _I, _VP, _VP, ctypes.c_double, _VP], _I),
    "tlbm_mlp2d_nu_t": ([_VP, _VP, _VP, _VP, _I, _I, _I, ctypes.c_double, _VP, _VP], _I),
    "tlbm_ai_les_bgk": ([_VP, _VP, _VP, _VP, _VP, _VP, _VP, _I, _I, ctypes.c_double, _VP], _I),
    "tlbm_mlp2d_grad": ([_VP, _VP, _VP, _VP, _I64, _VP, _VP], _I),
    "tlbm_mlp2d_adam": ([_VP, _VP, _VP, _I, _VP, _VP, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                         ctypes.c_float, ctypes.c_float, _VP, _VP], _I),
    "tlbm_coarsen_moments": ([_I, _I, _VP, _VP, _I, _I, _I, _I, _VP, _VP], _I),
    "tlbm_sgs_samples": ([_VP, _I, _I, _I, _VP, _VP, _VP], _I),
    "tlbm_cavities_layout": ([_VP, _I, _VP, _VP], _I),
    "tlbm_cavities_step": ([_I, _VP, _VP, _I64, _VP, _I, ctypes.c_double, _I, _VP], _I),
    "tlbm_cavities_graph": ([_I, _VP, _VP, _I64, _VP, _I, ctypes.c_double, _I, _I, _VP], _I),
    "tlbm_cavities_step2": ([_I, _VP, _VP, _I64, _VP, _I, ctypes.c_double, _I, _VP], _I),
    "tlbm_cavities_graph2": ([_I, _VP, _VP, _I64, _VP, _I, ctypes.c_double, _I, _I, _VP], _I),
    "tlbm_graph_launch": ([_VP, _VP], _I),
    "tlbm_graph_destroy": ([_VP], _I),
}

_lib: ctypes.CDLL | None = None


def library_path() -> Path:
    override = os.environ.get("TLBM_HIP_LIBRARY")
    return Path(override) if override else LIB_PATH


def lib() -> ctypes.CDLL:
    """Load (once) and return the HIP library; raise if it is not built."""
    global _lib
    if _lib is None:
        path = library_path()
        if not path.is_file():
            raise RuntimeError(
                f"{LIB_NAME} not found at {path}; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` "
                "(there is no CPU fallback)"
            )
        handle = ctypes.CDLL(str(path))
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _lib = handle
    return _lib


def exported_symbols() -> list[str]:
    return sorted(_SIGNATURES)


def check(rc: int, what: str) -> None:
    """Map a status code to the reference's exception types."""
    if rc == OK:
        return
    msg = lib().tlbm_last_error().decode(errors="replace") or lib().tlbm_status_string(rc).decode()
    text = f"{what}: {msg}"
    if rc in (ERR_INVALID, ERR_SHAPE, ERR_DTYPE, ERR_ALIAS):
        raise ValueError(text)
    if rc == ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    raise RuntimeError(text)


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


Strides = ctypes.c_int64 * 3


def grid_of(f: torch.Tensor) -> tuple[int, int, int, int]:
    """(Q, nz, ny, nx) of a population tensor (D2Q9 fields are (9, ny, nx))."""
    q = int(f.shape[0])
    if q == 9 and f.ndim == 3:
        return q, 1, int(f.shape[1]), int(f.shape[2])
    if q in (19, 27) and f.ndim == 4:
        return q, int(f.shape[1]), int(f.shape[2]), int(f.shape[3])
    raise ValueError(f"populations must be (9, ny, nx), (19, nz, ny, nx) or (27, nz, ny, nx); got {tuple(f.shape)}")


def strides_of(f: torch.Tensor) -> "Strides":
    """(q, z, y) element strides of a population tensor whose x stride is 1."""
    if f.stride(-1) != 1:
        raise ValueError("populations need a unit stride along x")
    if f.ndim == 3:
        return Strides(f.stride(0), 0, f.stride(1))
    return Strides(f.stride(0), f.stride(1), f.stride(2))


def stream_of(t: torch.Tensor) -> int:
    """Raw hipStream_t of torch's current stream on the tensor's device."""
    return torch.cuda.current_stream(t.device).cuda_stream


def require_gpu(t: torch.Tensor, what: str) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise RuntimeError(f"{what} must live on a GPU device (got {t.device}); the HIP path has no CPU fallback")


_STORE_OF_DTYPE = {
    torch.float32: STORE_F32,
    torch.float64: STORE_F64,
    torch.float16: STORE_F16,
    torch.bfloat16: STORE_BF16,
}


def store_code(dtype: torch.dtype, deviation: bool = False) -> int:
    try:
        code = _STORE_OF_DTYPE[dtype]
    except KeyError:
        raise ValueError(f"Unsupported dtype {dtype}; use float32, float64, float16 or bfloat16") from None
    if deviation:
        if code == STORE_F16:
            return STORE_F16_DEV
        if code == STORE_BF16:
            return STORE_BF16_DEV
        raise ValueError(f"deviation storage needs float16 or bfloat16, got {dtype}")
    return code


def make_collision(
    kind: int,
    tau: float,
    *,
    tau_mode: int = TAU_CONST,
    eq_rho_floor: bool = True,
    cs_delta_sq: float = 0.0,
    rates: tuple | list = (),
    tau_eff: torch.Tensor | None = None,
    mlp: int | None = None,
) -> Collision:
    c = Collision()
    c.kind = kind
    c.tau_mode = tau_mode
    c.eq_rho_floor = 1 if eq_rho_floor else 0
    c.tau = float(tau)
    c.cs_delta_sq = float(cs_delta_sq)
    for i, r in enumerate(rates):
        c.rate[i] = float(r)
    c.tau_eff = ptr(tau_eff)
    c.mlp = mlp
    return c


def make_boundary(kind: int = BC_NONE, u=(0.0, 0.0, 0.0), faces: int = 0) -> Boundary:
    b = Boundary()
    b.kind = kind
    b.faces = int(faces)
    for i in range(3):
        b.u[i] = float(u[i])
    return b
