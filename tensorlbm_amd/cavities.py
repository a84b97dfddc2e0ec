"""Many independent lid-driven cavities advanced by one launch per step (or one graph per N steps).

Each cavity ``i`` is a cubic D3Q19 BGK box of edge ``L_i`` with relaxation time ``τ_i``: the
reference's cavity chain (pull stream, Zou-He moving lid on y = L−1 from
``zou_he_moving_lid_3d``, boundaries3d.py:324-404, full-way bounce-back on the other five
faces, BGK) on its own grid.  ``BatchedCavities`` packs all of them into one device buffer
(cavity ``i`` at an element offset, ``[z][q][y][x]`` with a padded plane) and advances them with
``tlbm_cavities_step``; ``run(n, graph=True)`` replays a captured HIP graph of ping-pong steps.
With ``temporal_blocking=True`` (fp32 or 16-bit deviation storage) ``run`` advances two steps per
launch (``tlbm_cavities_step2``: the two-step march over 64 × 8 column tiles of every cavity, the
intermediate lattice kept on chip).

``autotune()`` picks the launch shape by timing each candidate on the actual batch: one workgroup
per 256-cell tile or a persistent grid of a given size for one step per launch, and, where the
storage allows it, the z chunk of the two-step launch; it switches to the fastest.
"""

from __future__ import annotations

import ctypes
import math
from typing import Sequence

import numpy as np
import torch

from . import _native as nat

__all__ = ["BatchedCavities", "cavity_batch_config", "wall_mask"]


def cavity_batch_config(n: int = 64, seed: int = 0, lo: int = 32, hi: int = 160, u_lid: float = 0.1,
                        re_range=(100.0, 1000.0)):
    """Seeded edges U{lo..hi}, Re log-uniform in re_range, τ = 3 u_lid L / Re + 0.5."""
    rng = np.random.default_rng(seed)
    edges = rng.integers(lo, hi + 1, size=n)
    re = np.exp(rng.uniform(math.log(re_range[0]), math.log(re_range[1]), size=n))
    taus = 3.0 * u_lid * edges / re + 0.5
    return [int(e) for e in edges], [float(t) for t in taus]


def wall_mask(edge: int, device=None) -> torch.Tensor:
    """(L, L, L) int8 mask of the five stationary walls: x = 0, x = L−1, y = 0, z = 0, z = L−1."""
    m = torch.zeros((edge, edge, edge), dtype=torch.int8, device=device)
    m[:, :, 0] = 1
    m[:, :, -1] = 1
    m[:, 0, :] = 1
    m[0] = 1
    m[-1] = 1
    return m


def _plane(edge: int) -> int:
    return (edge * edge + 63) // 64 * 64


class BatchedCavities:
    """A batch of cubic D3Q19 lid-driven cavities on one GPU."""

    def __init__(self, edges: Sequence[int], taus: Sequence[float], *, u_lid: float = 0.1,
                 device: str | torch.device = "cuda:0", dtype: torch.dtype = torch.float32,
                 deviation: bool = False, temporal_blocking: bool = False) -> None:
        if len(edges) != len(taus) or not 1 <= len(edges) <= nat.CAVITIES_MAX:
            raise ValueError(f"need 1..{nat.CAVITIES_MAX} cavities with one tau each")
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError(f"BatchedCavities requires a GPU device, got {self.device}")
        self.dtype, self.deviation = dtype, bool(deviation)
        self._store = nat.store_code(dtype, self.deviation)
        self.u_lid = float(u_lid)
        self.n = len(edges)
        self._table = (nat.Cavity * self.n)()
        for i, (e, t) in enumerate(zip(edges, taus)):
            self._table[i].edge = int(e)
            self._table[i].tau = float(t)
        tot, tiles = ctypes.c_int64(), ctypes.c_int64()
        nat.check(nat.lib().tlbm_cavities_layout(self._table, self.n, ctypes.byref(tot), ctypes.byref(tiles)),
                  "tlbm_cavities_layout")
        self.total_elements, self.total_tiles = tot.value, tiles.value
        self.edges = [int(e) for e in edges]
        self.taus = [float(t) for t in taus]
        # zero-filled: the plane padding is never written by the kernel
        self.a = torch.zeros(self.total_elements, dtype=dtype, device=self.device)
        self.b = torch.zeros_like(self.a)
        self.grid = 0
        # two steps per launch: fp32 or 16-bit deviation storage
        self.two_step_capable = self._store in (nat.STORE_F32, nat.STORE_F16_DEV, nat.STORE_BF16_DEV)
        if temporal_blocking and not self.two_step_capable:
            raise ValueError("two steps per launch take float32 or 16-bit deviation storage")
        self.two_step = bool(temporal_blocking)
        self.zc = 0  # z planes per two-step workgroup; 0: the launch-shape model picks
        self._graphs: dict = {}

    # -- layout --------------------------------------------------------------
    @property
    def num_cells(self) -> int:
        return sum(e**3 for e in self.edges)

    def offset(self, i: int) -> int:
        return int(self._table[i].offset)

    def cavity(self, i: int, buf: torch.Tensor | None = None) -> torch.Tensor:
        """(19, L, L, L) strided view of cavity i's populations in ``buf`` (default: current state)."""
        buf = self.a if buf is None else buf
        e = self.edges[i]
        p = _plane(e)
        return buf.as_strided((19, e, e, e), (p, 19 * p, e, 1), self.offset(i))

    def fill_equilibrium(self) -> None:
        """ρ = 1, u = 0 everywhere (stored as f − w for deviation storage)."""
        from .lattice import D3Q19

        w = D3Q19.W_EXACT64.to(self.device)
        for i in range(self.n):
            v = self.cavity(i)
            v.copy_((torch.zeros_like(w) if self.deviation else w).reshape(19, 1, 1, 1).expand_as(v))

    def populations(self, i: int) -> torch.Tensor:
        """Cavity i's populations f as a contiguous (19, L, L, L) tensor (w added back for deviation storage)."""
        v = self.cavity(i).to(torch.float64 if self.dtype == torch.float64 else torch.float32)
        if self.deviation:
            from .lattice import D3Q19

            v = v + D3Q19.W_EXACT64.to(device=self.device, dtype=v.dtype).reshape(19, 1, 1, 1)
        return v.contiguous()

    # -- stepping -------------------------------------------------------------
    def step(self) -> None:
        """One step of every cavity (state moves a -> b, then the buffers swap)."""
        nat.check(nat.lib().tlbm_cavities_step(self._store, self.a.data_ptr(), self.b.data_ptr(), self.a.numel(),
                                               self._table, self.n, self.u_lid, int(self.grid),
                                               nat.stream_of(self.a)), "tlbm_cavities_step")
        self.a, self.b = self.b, self.a

    def step2(self) -> None:
        """Two steps of every cavity: one launch with ``two_step``, else two."""
        if not self.two_step:
            self.step()
            self.step()
            return
        nat.check(nat.lib().tlbm_cavities_step2(self._store, self.a.data_ptr(), self.b.data_ptr(), self.a.numel(),
                                                self._table, self.n, self.u_lid, int(self.zc),
                                                nat.stream_of(self.a)), "tlbm_cavities_step2")
        self.a, self.b = self.b, self.a

    def _graph2(self, launches: int):
        key = ("step2", self.a.data_ptr(), self.b.data_ptr(), int(self.zc), int(launches))
        g = self._graphs.get(key)
        if g is None:
            h = ctypes.c_void_p()
            nat.check(nat.lib().tlbm_cavities_graph2(self._store, self.a.data_ptr(), self.b.data_ptr(),
                                                     self.a.numel(), self._table, self.n, self.u_lid,
                                                     int(self.zc), int(launches), ctypes.byref(h)),
                      "tlbm_cavities_graph2")
            g = h.value
            self._graphs[key] = g
        return g

    def _graph(self, steps: int):
        key = (self.a.data_ptr(), self.b.data_ptr(), int(self.grid), int(steps))
        g = self._graphs.get(key)
        if g is None:
            h = ctypes.c_void_p()
            nat.check(nat.lib().tlbm_cavities_graph(self._store, self.a.data_ptr(), self.b.data_ptr(),
                                                    self.a.numel(), self._table, self.n, self.u_lid,
                                                    int(self.grid), int(steps), ctypes.byref(h)),
                      "tlbm_cavities_graph")
            g = h.value
            self._graphs[key] = g
        return g

    def run(self, n_steps: int, *, graph: bool = True, graph_steps: int = 20) -> None:
        """Advance ``n_steps``: whole graphs of ``graph_steps`` ping-pong steps, eager launches for the rest.

        With ``two_step`` a graph holds ``graph_steps / 2`` two-step launches (``graph_steps`` a
        multiple of 4) and the eager rest runs in pairs, an odd last step with one step per launch.
        """
        n_steps = int(n_steps)
        if self.two_step:
            if graph and n_steps >= graph_steps:
                if graph_steps < 4 or graph_steps % 4:
                    raise ValueError("with two steps per launch graph_steps must be a multiple of 4")
                g = self._graph2(graph_steps // 2)
                s = nat.stream_of(self.a)
                for _ in range(n_steps // graph_steps):
                    nat.check(nat.lib().tlbm_graph_launch(g, s), "tlbm_graph_launch")
                n_steps %= graph_steps
            for _ in range(n_steps // 2):
                self.step2()
            if n_steps % 2:
                self.step()
            return
        if graph and n_steps >= graph_steps:
            if graph_steps < 2 or graph_steps % 2:
                raise ValueError("graph_steps must be even and >= 2")
            g = self._graph(graph_steps)
            s = nat.stream_of(self.a)
            for _ in range(n_steps // graph_steps):
                nat.check(nat.lib().tlbm_graph_launch(g, s), "tlbm_graph_launch")
            n_steps %= graph_steps
        for _ in range(n_steps):
            self.step()

    def close(self) -> None:
        for g in self._graphs.values():
            nat.lib().tlbm_graph_destroy(g)
        self._graphs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- tuning / timing --------------------------------------------------------
    def time_steps(self, n_steps: int = 20, *, graph: bool = True, graph_steps: int = 20) -> float:
        """Seconds per step over ``n_steps`` (HIP events on the launch stream)."""
        self.run(2, graph=False)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize(self.device)
        start.record()
        self.run(n_steps, graph=graph, graph_steps=graph_steps)
        end.record()
        end.synchronize()
        return start.elapsed_time(end) * 1e-3 / n_steps

    def autotune(self, candidates: Sequence[int] | None = None, n_steps: int = 10,
                 zc_candidates: Sequence[int] | None = None) -> dict:
        """Time every launch shape on this batch and keep the fastest; returns {shape: seconds per step}.

        Shapes: ``"grid=G"`` -- one step per launch, G = 0 one workgroup per tile, else a persistent
        grid of G workgroups (``candidates``); ``"zc=Z"`` -- two steps per launch with z chunks of Z
        planes, Z = 0 the launch-shape model's choice (``zc_candidates``; only where the storage
        allows two steps per launch).  Afterwards ``grid`` is the fastest one-step shape (used for
        odd steps) and ``two_step`` / ``zc`` follow the fastest shape overall.  The timed steps run on a
        copy of the populations: the state is restored, so a run tuned first stays on its own timeline.
        """
        saved = self.a.clone()
        cus = torch.cuda.get_device_properties(self.device).multi_processor_count
        grids = list(candidates) if candidates is not None else [0] + [cus * k for k in (4, 8, 16, 32)]
        zcs = list(zc_candidates) if zc_candidates is not None else [0, 8, 16, 32, 64]
        res, one = {}, {}
        self.two_step = False
        for g in grids:
            self.grid = int(g)
            one[int(g)] = res[f"grid={int(g)}"] = self.time_steps(n_steps, graph=False)
        self.grid = min(one, key=one.get)
        if self.two_step_capable:
            self.two_step = True
            for z in zcs:
                self.zc = int(z)
                res[f"zc={int(z)}"] = self.time_steps(n_steps, graph=False)
        best = min(res, key=res.get)
        self.two_step = best.startswith("zc=")
        if self.two_step:
            self.zc = int(best.split("=")[1])
        self.a.copy_(saved)
        del saved
        return res
