"""Compute/store precision tiers (mirror of src/tensorlbm/precision.py:72-158).

``<COMPUTE><STORE>`` names, e.g. ``FP32FP16`` = float32 arithmetic with
float16 storage.  On the HIP path the kernel is the cast boundary: loads widen
to the compute type in registers and stores narrow to the storage type.
"""

from __future__ import annotations

import enum

import torch

__all__ = ["PrecisionPolicy", "cast_to_compute", "cast_to_store"]


class PrecisionPolicy(enum.Enum):
    FP64FP64 = (torch.float64, torch.float64)
    FP64FP32 = (torch.float64, torch.float32)
    FP64FP16 = (torch.float64, torch.float16)
    FP32FP32 = (torch.float32, torch.float32)
    FP32FP16 = (torch.float32, torch.float16)

    def __init__(self, compute_dtype: torch.dtype, store_dtype: torch.dtype) -> None:
        self._compute_dtype = compute_dtype
        self._store_dtype = store_dtype

    @property
    def compute_dtype(self) -> torch.dtype:
        return self._compute_dtype

    @property
    def store_dtype(self) -> torch.dtype:
        return self._store_dtype

    def cast_to_compute(self, f: torch.Tensor) -> torch.Tensor:
        return cast_to_compute(f, self)

    def cast_to_store(self, f: torch.Tensor) -> torch.Tensor:
        return cast_to_store(f, self)

    @classmethod
    def parse(cls, value: "PrecisionPolicy | str | None") -> "PrecisionPolicy":
        """Coerce a policy, its exact name or None (-> FP32FP32); ValueError / TypeError otherwise."""
        if value is None:
            return cls.FP32FP32
        if isinstance(value, cls):
            return value
        if isinstance(value, str):
            try:
                return cls[value]
            except KeyError:
                valid = ", ".join(m.name for m in cls)
                raise ValueError(f"Unknown precision policy {value!r}; valid names: {valid}") from None
        raise TypeError(f"precision policy must be PrecisionPolicy, str, or None, got {type(value).__name__}")


def cast_to_compute(f: torch.Tensor, policy) -> torch.Tensor:
    policy = PrecisionPolicy.parse(policy)
    return f if f.dtype == policy.compute_dtype else f.to(policy.compute_dtype)


def cast_to_store(f: torch.Tensor, policy) -> torch.Tensor:
    policy = PrecisionPolicy.parse(policy)
    return f if f.dtype == policy.store_dtype else f.to(policy.store_dtype)
