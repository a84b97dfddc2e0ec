"""Throughput of the fused per-step lattice update on one MI355X.

Default workload (the headline D3Q19 256³ configuration): BGK, fp32,
256³ channel with a solid sphere (radius 24 cells, centre x=64, y=z=128) as a
full-way bounce-back mask, equilibrium inlet u=0.05 at x=0, zero-gradient
outlet, periodic in y/z, seeded 1e-3 velocity perturbation.  A step is the
fused pull-stream + inlet/outlet + bounce-back + collide update; the solver
advances two steps per launch (``k_lbm2``, intermediate lattice in LDS).

Prints ONE JSON line.  ``value`` is MLUPS (lattice cells updated per second,
1e6) with the populations resident in HBM; every form is timed as the median
of 5 windows of ``--steps`` steps (HIP events on the launch stream), after a
1 s clock warm-up and ``--warmup`` untimed steps.  ``roofline`` prices the
least bytes one launch must move (each population read and written once, plus
the mask byte) against the 8 TB/s HBM peak; ``step_equivalent_GBs`` is the
rate at 2·Q·itemsize bytes per cell-update, which a launch that keeps its
intermediate lattice on chip can exceed.  ``traffic`` is the HBM bytes one
launch actually moves (PMC).  ``forms`` reports the other storage/collision
forms of the same grid; ``cpu_baseline`` times the reference's CPU step chain
restated in eager torch on the host cores.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MLUPS & % HBM roofline, D3Q19 256³: plain FP32 vs FP16-store+MLP closure"
HBM_PEAK_GBS = 8000.0
SEED = 20260819


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=400)
    p.add_argument("--warmup", type=int, default=40)
    p.add_argument("--n", type=int, default=256, help="cubic grid edge")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-forms", action="store_true", help="only the headline form")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--copy-probe", action="store_true", help="also time a plain device copy of the same bytes")
    p.add_argument("--skip", default="", help="comma-separated forms to leave out: cavities, dhit")
    p.add_argument("--one-step", action="store_true", help="headline form with one step per launch (k_lbm)")
    p.add_argument("--layout", choices=("padded", "contiguous"), default="padded",
                   help="population layout of the headline form's solver buffers")
    return p.parse_args()


def channel_case(n: int, dev, torch):
    """Sphere mask and initial populations of the channel workload (built on the GPU)."""
    import tensorlbm_amd as tl

    g = torch.Generator(device="cpu").manual_seed(SEED)
    a = torch.arange(n, dtype=torch.float32, device=dev)
    z, y, x = torch.meshgrid(a, a, a, indexing="ij")
    r = 24.0 * n / 256.0
    solid = ((x - 64.0 * n / 256.0) ** 2 + (y - n / 2) ** 2 + (z - n / 2) ** 2) <= r * r
    del x, y, z
    pert = ((torch.rand((3, n, n, n), generator=g) * 2 - 1) * 1e-3).to(dev)
    ux = 0.05 + pert[0]
    ux[solid] = 0.0
    rho = torch.ones((n, n, n), device=dev)
    f = tl.equilibrium3d(rho, ux, pert[1], pert[2])
    del pert, ux, rho
    return solid.to(torch.int8).contiguous(), f


REPS = 5  # timed windows per form; the reported time is their median


def _median(v):
    v = sorted(v)
    k = len(v)
    return v[k // 2] if k % 2 else 0.5 * (v[k // 2 - 1] + v[k // 2])


def warm_clocks(step, f, seconds, torch):
    """Launch the step for about ``seconds`` so the GPU leaves its idle clocks before any timing."""
    t0 = time.perf_counter()
    n = 0
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            f = step(f)
        torch.cuda.synchronize()
        n += 10
    return f, n


def time_steps(step, f, steps, warmup, torch, reps=REPS):
    """Median over ``reps`` windows of ``steps`` launches of the ms per launch, HIP events on torch's
    current stream (the launch stream), after ``warmup`` untimed launches."""
    for _ in range(warmup):
        f = step(f)
    torch.cuda.synchronize()
    windows = []
    t0 = time.perf_counter()
    for _ in range(reps):
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(steps):
            f = step(f)
        stop.record()
        torch.cuda.synchronize()
        windows.append(start.elapsed_time(stop) / steps)
    wall = time.perf_counter() - t0
    return _median(windows), {"min": min(windows), "max": max(windows), "reps": reps, "wall_s": wall}, f


def per_launch_times(step, f, launches, torch):
    """ms of each of ``launches`` launches, one event pair each (the spread between launches)."""
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(launches)]
    torch.cuda.synchronize()
    for a, b in evs:
        a.record()
        f = step(f)
        b.record()
    torch.cuda.synchronize()
    return [a.elapsed_time(b) for a, b in evs], f


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(n: int, tau: float, seconds: float):
    """The reference's CPU path, restated in eager torch (collide_bgk3d -> stream3d_roll -> channel
    inlet/outlet -> obstacle bounce-back), timed on the full grid for a bounded number of steps."""
    import numpy as np
    import torch

    from oracle import ref, torch_cpu

    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    rng = np.random.default_rng(SEED)
    solid = torch.from_numpy(ref.sphere_mask(n, n, n, 64.0 * n / 256.0, n / 2, n / 2, 24.0 * n / 256.0))
    pert = torch.from_numpy(rng.uniform(-1e-3, 1e-3, (3, n, n, n)).astype(np.float32))
    ux = 0.05 + pert[0]
    ux[solid] = 0.0
    f = torch_cpu.equilibrium3d(torch.ones((n, n, n)), ux, pert[1], pert[2])
    del pert, ux
    f = torch_cpu.channel_step(f, tau, 0.05, solid)  # warm-up (allocator, page faults)
    steps, t0 = 0, time.perf_counter()
    times = []
    while True:
        t1 = time.perf_counter()
        f = torch_cpu.channel_step(f, tau, 0.05, solid)
        times.append(time.perf_counter() - t1)
        steps += 1
        el = time.perf_counter() - t0
        if el > seconds or steps >= 200:
            break
    ms = _median(times) * 1e3
    finite = bool(torch.isfinite(f[:, ::8, ::8, ::8]).all())
    del f
    return {
        "value": round(n**3 / (ms * 1e-3) / 1e6, 2),
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} eager-torch steps (median {ms:.0f} ms) of the reference's CPU chain collide_bgk3d -> "
                  f"stream3d_roll -> channel inlet/outlet -> obstacle bounce-back on the full {n}^3 fp32 channel + "
                  f"sphere grid, {el:.1f} s on {threads} torch threads",
        "cpu_model": cpu_model(),
        "torch": torch.__version__,
        "finite": finite,
    }


PMC_SUMMARY = {1: ROOT / "profiles" / "k_lbm_d3q19_bgk_fp32_channel256_pmc.md",
               2: ROOT / "profiles" / "k_lbm2_d3q19_bgk_fp32_channel256_pmc.md"}


def measured_traffic(n: int, steps_per_launch: int):
    """HBM bytes per launch of the headline kernel from the committed PMC summary (256^3 only)."""
    path = PMC_SUMMARY[steps_per_launch]
    if n != 256 or not path.is_file():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":", 1)[1]))
    return None


def dhit_forms(torch, dev, n, steps=100):
    """Decaying isotropic turbulence, D3Q27 Geier cumulants, τ = 0.505, random solenoidal start
    (k0 = 4, u_rms = 0.03): throughput of fp32 and of the fp16 / bf16 deviation storages."""
    import tensorlbm_amd as tl
    from tensorlbm_amd import fields

    u = fields.random_solenoidal_field(n, 4.0, 0.03, seed=SEED, device=dev, dtype=torch.float32)
    f0 = tl.equilibrium27(torch.ones((n, n, n), device=dev), u[0].contiguous(), u[1].contiguous(), u[2].contiguous())
    del u
    w = tl.lattice.D3Q27.W_EXACT64.to(dev).view(27, 1, 1, 1)
    out = {}
    for name, dtype, dv in (("fp32", torch.float32, False), ("fp16dev", torch.float16, True),
                            ("bf16dev", torch.bfloat16, True)):
        f = (f0 - w).to(dtype) if dv else f0.clone()
        s = tl.FusedSolver3D(n, n, n, 0.505, dev, dtype, lattice="D3Q27", collision="GEIER", deviation=dv)
        per = 2 if s.two_step else 1
        ms_launch, _, f = time_steps(s.step2 if per == 2 else s.step, f, steps // per, 10, torch)
        ms = ms_launch / per
        bpc = 2 * 27 * torch.tensor([], dtype=dtype).element_size()
        out[name] = {"ms_per_step": round(ms, 5), "mlups": round(n**3 / (ms * 1e-3) / 1e6, 1), "bytes_per_cell": bpc,
                     "steps_per_launch": per,
                     "hbm_frac": round(bpc * n**3 / (ms_launch * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        del s, f
    del f0
    torch.cuda.empty_cache()
    return out


def cavities_form(torch, dev, steps=100):
    """64 lid-driven cavities (edges U{32..160}, Re log-uniform [100, 1000]): one autotuned launch per step
    (or per two steps, when the autotuner picks the two-step launch) replayed from a HIP graph, against 64
    single-cavity fused launches per step."""
    import tensorlbm_amd as tl
    from tensorlbm_amd.cavities import BatchedCavities, cavity_batch_config, wall_mask

    edges, taus = cavity_batch_config(64, seed=SEED)
    b = BatchedCavities(edges, taus, u_lid=0.1, device=dev, dtype=torch.float32)
    b.fill_equilibrium()
    tune = b.autotune()
    two_step = b.two_step
    ms_graph = b.time_steps(steps, graph=True, graph_steps=20) * 1e3
    ms_eager = b.time_steps(steps // 2, graph=False) * 1e3
    cells = b.num_cells
    ok = bool(torch.isfinite(b.populations(0)).all())
    b.close()
    del b
    solvers = []
    for e, t in zip(edges, taus):
        s = tl.FusedSolver3D(e, e, e, t, dev, torch.float32, solid=wall_mask(e, dev), boundary="lid_y",
                             u_bc=(0.1, 0.0, 0.0))
        f = tl.equilibrium3d(torch.ones((e, e, e), device=dev), *(torch.zeros((e, e, e), device=dev),) * 3)
        solvers.append([s, f])

    def seq_step(_):
        for item in solvers:
            item[1] = item[0].step(item[1])
        return None

    ms_seq, _, _ = time_steps(seq_step, None, steps // 4, 3, torch)
    del solvers
    torch.cuda.empty_cache()
    return {"cells": cells, "edges_min_max": [min(edges), max(edges)], "shape": tune and min(tune, key=tune.get),
            "steps_per_launch": 2 if two_step else 1,
            "autotune_ms": {str(k): round(v * 1e3, 4) for k, v in tune.items()},
            "ms_per_step_graph": round(ms_graph, 4), "mlups_graph": round(cells / (ms_graph * 1e-3) / 1e6, 1),
            "ms_per_step_eager": round(ms_eager, 4), "ms_per_step_64_sequential": round(ms_seq, 4),
            "mlups_64_sequential": round(cells / (ms_seq * 1e-3) / 1e6, 1),
            # least bytes per launch: every population read and written once (no mask: walls from coordinates)
            "hbm_frac_graph": round(cells * 152 / (ms_graph * (2 if two_step else 1) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "step_equivalent_GBs_graph": round(cells * 152 / (ms_graph * 1e-3) / 1e9, 1), "finite": ok}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the per-step update runs on one GPU; a z-slab split with one halo exchange per step is the "
              "multi-GPU design, not built)", file=sys.stderr)
        sys.exit(2)
    import torch

    import tensorlbm_amd as tl

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    n = args.n
    cells = n**3
    u_in = 0.05
    tau = 0.572  # Re = u·D/ν = 0.05·48/0.024 = 100
    solid, f0 = channel_case(n, dev, torch)
    q = 19

    closure = None

    def form(dtype, collision="BGK", deviation=False, boundary=True, mask=True, mlp=False, two_step=None,
             spread=False, force=False, layout="padded"):
        f = f0.to(dtype)
        if deviation:
            f = (f0 - tl.lattice.D3Q19.W_EXACT64.to(dev).view(19, 1, 1, 1)).to(dtype)
        solver = tl.FusedSolver3D(n, n, n, tau=tau, device=dev, dtype=dtype, collision=collision,
                                  deviation=deviation, solid=solid if mask else None,
                                  boundary="channel_x" if boundary else None, u_bc=(u_in, 0.0, 0.0),
                                  closure=closure if mlp else None, C_s=0.17 if mlp else 0.0,
                                  temporal_blocking=two_step, layout=layout)
        per = 2 if solver.two_step else 1
        launch = solver.step2 if per == 2 else solver.step
        acc = None
        if force:  # Ladd force of every step added to one FP64 accumulator (the caller zeroes it once)
            acc = torch.zeros(3, dtype=torch.float64, device=dev)
            launch = (lambda g: solver.step2(g, force=acc)) if per == 2 else (lambda g: solver.step(g, force=acc))
        if not warmed:
            f, _ = warm_clocks(launch, f, 1.0, torch)
            warmed.append(True)
        ms_launch, win, f = time_steps(launch, f, max(1, args.steps // per), max(1, args.warmup // per), torch)
        ms = ms_launch / per
        esize = torch.tensor([], dtype=dtype).element_size()
        bpc = 2 * q * esize
        # least bytes one launch must move: every population read and written once, plus the mask byte
        least = (bpc + (1 if mask else 0)) * cells
        res = {"ms_per_step": round(ms, 5), "mlups": round(cells / (ms * 1e-3) / 1e6, 1),
               "steps_per_launch": per, "ms_per_launch": round(ms_launch, 5),
               "ms_per_launch_window_min_max": [round(win["min"], 5), round(win["max"], 5)],
               "bytes_per_cell_update": bpc, "least_bytes_per_launch": least,
               "achieved_GBs": round(least / (ms_launch * 1e-3) / 1e9, 1),
               "hbm_frac": round(least / (ms_launch * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
               "step_equivalent_GBs": round(bpc * cells / (ms * 1e-3) / 1e9, 1),
               "windows": win["reps"], "wall_s": round(win["wall_s"], 3)}
        if spread:
            each, f = per_launch_times(launch, f, max(5, args.steps // per), torch)
            res["ms_per_launch_each_min_median_max"] = [round(min(each), 5), round(_median(each), 5),
                                                       round(max(each), 5)]
        res["finite"] = bool(torch.isfinite(f[:, ::16, ::16, ::16].float()).all())
        if acc is not None:
            res["force_finite"] = bool(torch.isfinite(acc).all())
        del solver, f
        return res

    warmed = []
    head = form(torch.float32, two_step=False if args.one_step else None, spread=True, layout=args.layout)
    forms = {}
    if not args.no_forms:
        forms["bgk_fp32_channel_sphere_one_step_per_launch"] = form(torch.float32, two_step=False)
        forms["bgk_fp32_channel_sphere_ladd_force"] = form(torch.float32, force=True)
        forms["bgk_fp32_periodic_nomask"] = form(torch.float32, boundary=False, mask=False)
        forms["mrt_fp32_channel_sphere"] = form(torch.float32, collision="MRT")
        forms["bgk_fp16dev_channel_sphere"] = form(torch.float16, deviation=True)
        forms["bgk_bf16dev_channel_sphere"] = form(torch.bfloat16, deviation=True)
        forms["bgk_fp16dev_periodic_nomask"] = form(torch.float16, deviation=True, boundary=False, mask=False)
        # learned closure evaluated in the collision (7 -> 32 -> 32 -> 1 on the matrix cores); the weights
        # are seeded (throughput does not depend on their values), the inputs normalised to the strain scale
        from tensorlbm_amd.closure import EddyViscosityMLP3D

        closure = EddyViscosityMLP3D(dev, seed=SEED)
        closure.set_feature_stats(torch.zeros(7), torch.full((7,), 1e-3))
        forms["bgk_fp16dev_mlp_channel_sphere"] = form(torch.float16, deviation=True, mlp=True)
        forms["bgk_fp32_mlp_channel_sphere"] = form(torch.float32, mlp=True)
        # the same network with bf16 matrix-core operands ("BF16 weights, FP32 accumulation")
        closure = EddyViscosityMLP3D(dev, params=closure.params, operands="bf16")
        forms["bgk_fp16dev_mlp_bf16_channel_sphere"] = form(torch.float16, deviation=True, mlp=True)
        # the closure form against the headline (fp32, two steps per launch) and against the plain
        # one-step fp32 kernel
        mlp_rate = forms["bgk_fp16dev_mlp_channel_sphere"]["mlups"]
        forms["fp16dev_mlp_speedup_vs_fp32"] = round(mlp_rate / head["mlups"], 3)
        forms["fp16dev_mlp_speedup_vs_fp32_one_step"] = round(
            mlp_rate / forms["bgk_fp32_channel_sphere_one_step_per_launch"]["mlups"], 3)
        skip = {x.strip() for x in args.skip.split(",") if x.strip()}
        if "cavities" not in skip:
            forms["cavities_64"] = cavities_form(torch, dev)
        if "dhit" not in skip:
            forms["dhit_d3q27_geier"] = dhit_forms(torch, dev, n)
    if args.copy_probe:
        src = f0.clone()
        dst = torch.empty_like(src)
        ms, _, _ = time_steps(lambda t: dst.copy_(src), src, args.steps, args.warmup, torch)
        forms["device_copy_probe"] = {"ms": round(ms, 5), "bytes": 2 * src.numel() * 4,
                                      "GBs": round(2 * src.numel() * 4 / (ms * 1e-3) / 1e9, 1)}
        del src, dst
    del f0
    torch.cuda.empty_cache()

    ms = head["ms_per_step"]
    value = head["mlups"]
    per = head["steps_per_launch"]
    kernel = ("k_lbm2<D3Q19, fp32, BGK, mask + channel, 64x8 tile> (two steps per launch)" if per == 2 else
              "k_lbm<D3Q19, fp32, BGK, stream, mask + boundaries>")
    traffic = measured_traffic(n, per)
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"D3Q19 BGK fp32 {n}^3 channel, sphere r=24 at x=64 (full-way bounce-back), "
                        "equilibrium inlet u=0.05 / copy outlet in x, periodic y/z, seeded 1e-3 perturbation",
            "grid": [n, n, n],
            "tau": tau,
            "lattice": "D3Q19",
            "collision": "BGK",
            "storage": "fp32",
            "layout": args.layout,
            "solid_cells": int(solid.sum().item()),
            "seed": SEED,
            "timing": f"median of {REPS} windows of {args.steps} steps after {args.warmup} warm-up steps and a "
                      "1 s clock warm-up; HIP events on the launch stream",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": head["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": head["hbm_frac"],
            "traffic": traffic,
            "traffic_GBs": round(traffic / (head["ms_per_launch"] * 1e-3) / 1e9, 1) if traffic else None,
            "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated; " + str(
                PMC_SUMMARY[per].relative_to(ROOT)),
            "least_bytes_per_launch": head["least_bytes_per_launch"],
            "least_bytes": "each population read and written once per launch (2 x 19 x 4 B per cell) plus the "
                           "1-byte solid mask",
            "step_equivalent_GBs": head["step_equivalent_GBs"],
            "bytes_per_cell_update": head["bytes_per_cell_update"],
            "steps_per_launch": per,
            "ms_per_launch": head["ms_per_launch"],
            "ms_per_launch_each_min_median_max": head["ms_per_launch_each_min_median_max"],
            "kernel": kernel,
        },
        "forms": forms,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(n, tau, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
