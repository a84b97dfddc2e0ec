"""Compile ``libtlbm_hip.so`` for gfx950 with hipcc (in-tree, no JIT cache)."""

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


SOURCES = ("tlbm.hip", "closure.hip", "cavities.hip")
OBJ_DIR = PKG / "_build"


def build(verbose: bool = False) -> Path:
    """Compile every source to an object in parallel, then link the shared library."""
    OBJ_DIR.mkdir(exist_ok=True)
    common = [hipcc(), *HIPCC_FLAGS, f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    jobs = []
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
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
