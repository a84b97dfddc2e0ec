"""Compile ``libtlbm_hip.so`` for gfx950 with hipcc and the ``torch.ops.tlbm`` library over it
(``libtlbm_torch.so``, host C++ against torch's headers); in-tree, no JIT cache."""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

CSRC = Path(__file__).resolve().parent
PKG = CSRC.parent
ROOT = PKG.parent
OUT = PKG / "libtlbm_hip.so"

HIPCC_FLAGS = [
    "-O3",
    "--offload-arch=gfx950",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-mcode-object-version=5",
    # no packed fp32 (v_pk_*_f32) instructions anywhere in the library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


SOURCES = ("tlbm.hip", "closure.hip", "closure2d.hip", "cavities.hip")
OBJ_DIR = PKG / "_build"
TORCH_OPS_SRC = CSRC / "torch_ops.cpp"
TORCH_OPS_OUT = PKG / "libtlbm_torch.so"


def _torch_flags() -> tuple[list[str], list[str]]:
    """Compile and link flags of the torch.ops.tlbm library (host C++ against torch's headers)."""
    import torch
    from torch.utils import cpp_extension as ce

    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cflags = ["-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
              f"-D_GLIBCXX_USE_CXX11_ABI={int(torch._C._GLIBCXX_USE_CXX11_ABI)}", f"-I{rocm}/include",
              f"-I{ROOT / 'include'}", *(f"-I{p}" for p in ce.include_paths())]
    tlib = ce.library_paths()[0]
    lflags = ["-shared", f"-L{tlib}", f"-Wl,-rpath,{tlib}", f"-L{PKG}", "-Wl,-rpath,$ORIGIN", f"-L{rocm}/lib",
              "-ltlbm_hip", "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-lamdhip64"]
    return cflags, lflags


def build(verbose: bool = False) -> Path:
    """Compile every source to an object in parallel, then link the shared library."""
    OBJ_DIR.mkdir(exist_ok=True)
    common = [hipcc(), *HIPCC_FLAGS, f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    jobs = []
    tcflags, tlflags = _torch_flags()
    tobj = OBJ_DIR / "torch_ops.o"
    tcmd = ["g++", *tcflags, "-c", "-o", str(tobj), str(TORCH_OPS_SRC)]
    if verbose:
        print(" ".join(tcmd), flush=True)
    tproc = subprocess.Popen(tcmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    for src in SOURCES:
        obj = OBJ_DIR / (Path(src).stem + ".o")
        cmd = [c for c in common if c != "-shared"] + ["-c", "-o", str(obj), str(CSRC / src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    failed = []
    for src, _, proc in jobs:
        out, err = proc.communicate()
        if proc.returncode != 0:
            sys.stderr.write(out + err)
            failed.append(src)
    if failed:
        raise RuntimeError(f"hipcc failed for {', '.join(failed)}")
    tmp = OUT.with_suffix(".so.tmp")
    link = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *[str(o) for _, o, _ in jobs]]
    if verbose:
        print(" ".join(link), flush=True)
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"link failed ({res.returncode})")
    os.replace(tmp, OUT)
    out, err = tproc.communicate()
    if tproc.returncode != 0:
        sys.stderr.write(out + err)
        raise RuntimeError("g++ failed for torch_ops.cpp")
    ttmp = TORCH_OPS_OUT.with_suffix(".so.tmp")
    tlink = ["g++", str(tobj), "-o", str(ttmp), *tlflags]
    if verbose:
        print(" ".join(tlink), flush=True)
    res = subprocess.run(tlink, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"link of libtlbm_torch.so failed ({res.returncode})")
    os.replace(ttmp, TORCH_OPS_OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
