"""DNS → learned closure → LES, end to end on one GPU (the 3-D ``run_ai_les_pipeline``).

The reference's pipeline (src/tensorlbm/ai/pipeline.py:222-398) runs a finer BGK reference,
block-averages it, trains ``EddyViscosityMLP`` and validates an LES that uses it.  Here:

1. ``run_dns_reference_3d``: fine D3Q19 BGK from the random solenoidal field, filtered by
   ``dns_scale``³ blocks every ``sample_every`` coarse steps (training run, one seed);
2. ``sgs_samples`` + ``train_eddy_viscosity_model_3d`` on the GPU kernels;
3. a second DNS (another seed) is the truth; LES runs of the same filtered initial state —
   no closure, Smagorinsky, learned closure — in the storage given by ``les_dtype``
   (``float16`` uses the deviation form), closure evaluated inside the fused step;
4. per snapshot: velocity rel-L2 and kinetic-energy error against the filtered DNS.

``python -m tensorlbm_amd.les_pipeline --n-les 192 --dns-scale 3`` is the 576³ → 192³ case
(about 29 GB for the DNS); smaller sizes run in seconds.  ``--k0`` moves the spectrum peak of
the initial field; training uses at most ``max_samples`` (a seeded random subset).
"""

from __future__ import annotations

import argparse
import json
import time

import torch

from . import closure as cl
from .fused import FusedSolver3D
from .lattice import D3Q19
from .ops import macroscopic3d

__all__ = ["run_ai_les_pipeline_3d"]


def _les(test, tau, steps, every, *, closure=None, C_s=0.0, dtype=torch.float32, device="cuda:0", keep_final=False):
    n = int(test[0].shape[1])
    f = cl.les_initial_populations(test[0])
    dev_store = dtype in (torch.float16, torch.bfloat16)
    if dev_store:
        f = (f - D3Q19.W_EXACT64.to(f.device, f.dtype).reshape(19, 1, 1, 1)).to(dtype)
    else:
        f = f.to(dtype)
    s = FusedSolver3D(n, n, n, tau, device, dtype, deviation=dev_store, closure=closure, C_s=C_s)
    errs, kes = [], []
    t0 = time.perf_counter()
    for k in range(every, steps + 1, every):
        f, _ = s.run(f, every)
        g = f.float()
        if dev_store:
            g = g + D3Q19.W_EXACT64.to(g.device, g.dtype).reshape(19, 1, 1, 1)
        _, ux, uy, uz = macroscopic3d(g.contiguous())
        u = torch.stack([ux, uy, uz]).double()
        ref = test[k // every][1:4].double()
        errs.append(float(torch.linalg.vector_norm(u - ref) / torch.linalg.vector_norm(ref)))
        kes.append(float((u * u).sum() / (ref * ref).sum()) - 1.0)
    torch.cuda.synchronize()
    out = {"err_final": errs[-1], "err_mean": sum(errs) / len(errs), "ke_err_final": kes[-1],
           "ke_err_mean_abs": sum(abs(v) for v in kes) / len(kes), "wall_s": time.perf_counter() - t0}
    if keep_final:
        out["u_final"] = u
    return out


def run_ai_les_pipeline_3d(n_les: int = 32, dns_scale: int = 3, tau_dns: float = 0.51, steps: int = 300,
                           sample_every: int = 10, train_seed: int = 101, test_seed: int = 202, epochs: int = 30,
                           C_s: float = 0.17, les_dtype: torch.dtype = torch.float32,
                           device: str | torch.device = "cuda:0", log=None, k0: float = 4.0,
                           u_rms: float = 0.03, max_samples: int = 1 << 25, precision_check: bool = False,
                           closure_operands: str = "fp16") -> dict:
    """Train on one DNS, validate LES runs against another; see the module docstring.

    ``precision_check`` also runs the learned-closure LES in fp64 storage and reports the
    velocity rel-L2 of the ``les_dtype`` run against it after the last step
    (``learned_vs_fp64_rel_l2``).  ``closure_operands`` ("fp16" or "bf16") picks the matrix-core
    form the fused collision evaluates the trained network in.
    """
    def say(msg):
        if log is not None:
            log(f"[les_pipeline {time.perf_counter() - t0:7.1f} s] {msg}")

    t0 = time.perf_counter()
    train = cl.run_dns_reference_3d(n_les, tau_dns, steps, sample_every, train_seed, device, dns_scale=dns_scale,
                                    k0=k0, u_rms=u_rms)
    x, t = zip(*(cl.sgs_samples(s) for s in train[1:]))
    del train
    feats, labels = torch.cat(x), torch.cat(t)
    del x, t
    say(f"training DNS done: {feats.shape[0]} samples")
    if feats.shape[0] > max_samples:  # a seeded random subset bounds the training time
        keep = torch.randperm(feats.shape[0], generator=torch.Generator().manual_seed(train_seed))[:max_samples]
        keep = keep.to(feats.device)
        feats, labels = feats.index_select(0, keep), labels.index_select(0, keep)
        say(f"kept {max_samples} of them")
    t_dns = time.perf_counter() - t0
    model, meta = cl.train_eddy_viscosity_model_3d(feats, labels, cl.TrainConfig3D(epochs=epochs), device=device,
                                                   log=say)
    del feats, labels
    if closure_operands != "fp16":
        model = cl.EddyViscosityMLP3D(model.device, params=model.params, operands=closure_operands)
    say(f"trained: val R2 {meta['final_val_r2']:.3f} in {meta['training_time_s']:.1f} s")
    test = cl.run_dns_reference_3d(n_les, tau_dns, steps, sample_every, test_seed, device, dns_scale=dns_scale,
                                   k0=k0, u_rms=u_rms)
    tau = cl.les_tau(tau_dns, dns_scale)
    say("test DNS done")
    runs = {}
    for name, kw in (("no_closure", {}), ("smagorinsky", {"C_s": C_s}), ("learned", {"closure": model, "C_s": C_s})):
        keep = precision_check and name == "learned"
        runs[name] = _les(test, tau, steps, sample_every, dtype=les_dtype, device=device, keep_final=keep, **kw)
        say(f"LES {name}: { {k: v for k, v in runs[name].items() if k != 'u_final'} }")
    res = {"n_les": n_les, "n_dns": n_les * dns_scale, "tau_dns": tau_dns, "tau_les": tau, "steps": steps,
           "k0": k0, "u_rms": u_rms,
           "les_dtype": str(les_dtype).replace("torch.", ""), "closure_operands": closure_operands,
           "samples_train": meta["n_samples_train"],
           "val_r2": meta["final_val_r2"], "train_s": meta["training_time_s"], "dns_s": t_dns, "les": runs}
    if precision_check:
        u_low = runs["learned"].pop("u_final")
        ref64 = _les(test, tau, steps, sample_every, dtype=torch.float64, device=device, keep_final=True,
                     closure=model, C_s=C_s)
        u64 = ref64.pop("u_final")
        res["learned_fp64"] = ref64
        res["learned_vs_fp64_rel_l2"] = float(torch.linalg.vector_norm(u_low - u64) / torch.linalg.vector_norm(u64))
        say(f"learned closure, {res['les_dtype']} vs fp64 storage: velocity rel-L2 {res['learned_vs_fp64_rel_l2']:.3e}")
        del u_low, u64
    return res


def main():
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    p.add_argument("--n-les", type=int, default=32)
    p.add_argument("--dns-scale", type=int, default=3)
    p.add_argument("--tau-dns", type=float, default=0.51)
    p.add_argument("--steps", type=int, default=300)
    p.add_argument("--every", type=int, default=10)
    p.add_argument("--epochs", type=int, default=30)
    p.add_argument("--les-dtype", choices=("float32", "float16"), default="float32")
    p.add_argument("--k0", type=float, default=4.0, help="spectrum peak of the initial field (box wavenumbers)")
    p.add_argument("--u-rms", type=float, default=0.03)
    p.add_argument("--max-samples", type=int, default=1 << 25)
    p.add_argument("--closure-operands", choices=("fp16", "bf16"), default="fp16")
    p.add_argument("--precision-check", action="store_true",
                   help="also run the learned-closure LES in fp64 storage and compare")
    a = p.parse_args()
    res = run_ai_les_pipeline_3d(a.n_les, a.dns_scale, a.tau_dns, a.steps, a.every, epochs=a.epochs,
                                 les_dtype=getattr(torch, a.les_dtype), log=lambda m: print(m, flush=True),
                                 k0=a.k0, u_rms=a.u_rms, max_samples=a.max_samples,
                                 precision_check=a.precision_check, closure_operands=a.closure_operands)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
