"""Step-context and reporter hooks for the host loops (mirror of src/tensorlbm/reporters.py:88-370).

A reporter is any callable taking a :class:`StepContext`; it may carry an
``interval`` attribute (fire every ``interval`` completed steps).
:class:`ThroughputReporter` measures MLUPS between firings after a device
synchronisation, as the reference's does.
"""

from __future__ import annotations

from time import perf_counter
from typing import Any, Callable, Sequence, TextIO

import torch


class StepContext:
    __slots__ = ("step", "f", "diag", "lattice", "num_cells", "num_steps", "macroscopic_fn", "units",
                 "state", "stop")

    def __init__(self, step: int = 0, f: torch.Tensor | None = None, diag: dict | None = None,
                 lattice: str = "D3Q19", num_cells: int = 0, num_steps: int = 0,
                 macroscopic_fn: Callable | None = None, units: Any = None) -> None:
        self.step = step
        self.f = f
        self.diag = diag if diag is not None else {}
        self.lattice = lattice
        self.num_cells = num_cells
        self.num_steps = num_steps
        self.macroscopic_fn = macroscopic_fn
        self.units = units
        self.state: dict = {}
        self.stop = False

    def macroscopic(self):
        if self.macroscopic_fn is not None:
            return self.macroscopic_fn(self.f)
        from . import ops

        q = int(self.f.shape[0])
        return {9: ops.macroscopic, 19: ops.macroscopic3d, 27: ops.macroscopic27}[q](self.f)


class ReporterBase:
    def __init__(self, interval: int = 1) -> None:
        if int(interval) < 1:
            raise ValueError(f"interval must be >= 1, got {interval}")
        self.interval = int(interval)


class CallbackReporter(ReporterBase):
    def __init__(self, callback: Callable[[StepContext], None], interval: int = 1, name: str = "callback"):
        super().__init__(interval)
        self.callback = callback
        self.name = name
        self.calls = 0

    def __call__(self, ctx: StepContext) -> None:
        self.calls += 1
        self.callback(ctx)


class ThroughputReporter(ReporterBase):
    """MLUPS = lattice updates / 1e6 / elapsed s between consecutive firings (reporters.py:289-370)."""

    def __init__(self, interval: int = 1, *, num_cells: int | None = None, out: TextIO | None = None,
                 sync: bool = True) -> None:
        super().__init__(interval)
        self.num_cells = num_cells
        self.out = out
        self.sync = bool(sync)
        self.records: list[tuple[int, float, float]] = []
        self._t_last: float | None = None
        self._step_last: int | None = None

    @property
    def last_mlups(self):
        return self.records[-1][2] if self.records else None

    @property
    def mean_mlups(self):
        return sum(r[2] for r in self.records) / len(self.records) if self.records else None

    def __call__(self, ctx: StepContext) -> None:
        if self.sync and ctx.f is not None and ctx.f.device.type == "cuda":
            torch.cuda.synchronize(ctx.f.device)
        now = perf_counter()
        n = self.num_cells if self.num_cells is not None else ctx.num_cells
        if self._t_last is None:
            self._t_last, self._step_last = now, ctx.step
            return
        elapsed, steps = now - self._t_last, ctx.step - self._step_last
        self._t_last, self._step_last = now, ctx.step
        if elapsed <= 0.0 or steps <= 0:
            return
        mlups = steps * n / 1e6 / elapsed
        self.records.append((ctx.step, elapsed, mlups))
        if self.out is not None:
            print(f"{ctx.step} {elapsed:.3f} {mlups:.1f}", file=self.out)


def dispatch(ctx: StepContext, reporters: Sequence[Callable]) -> None:
    """Call every reporter whose interval divides the completed step count."""
    for r in reporters:
        interval = int(getattr(r, "interval", 1))
        if ctx.step % interval == 0:
            r(ctx)
