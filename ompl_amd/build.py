"""Builds libompl_amd.so (HIP kernels + C-ABI) for gfx950, in-tree.

Usage: python -m ompl_amd.build [--jobs N] [--debug]

Every source under ompl_amd/csrc is compiled by hipcc with
  --offload-arch=gfx950 -O3 -ffp-contract=off
  -Xclang -target-feature -Xclang -packed-fp32-ops
(-ffp-contract=off keeps every FP32/FP64 expression in source order, the
basis of bit-exact parity; packed FP32 VALU ops are disabled) and linked into
ompl_amd/libompl_amd.so next to this file.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libompl_amd.so"
ROOT = PKG.parent

ARCH = "gfx950"
COMMON = [
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Wall",
    "-Wno-unused-function",
    "-I" + str(ROOT / "include"),
    "-I" + str(CSRC),
]
DEVICE = [
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: libompl_amd needs ROCm's hipcc to build")


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path, opt: list[str]) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [hipcc(), "-c", str(src), "-o", str(obj)] + COMMON + opt
    if src.suffix == ".hip":
        cmd += DEVICE + ["-x", "hip"]
    else:
        cmd += DEVICE
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, debug: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    opt = ["-O1", "-g"] if debug else ["-O3"]
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, opt), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-o", str(tmp)] + [str(o) for o in objs] + [
        "--offload-arch=" + ARCH, "-fPIC", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--debug", action="store_true")
    a = ap.parse_args(argv)
    lib = build(a.jobs, a.debug)
    print(lib)
    return 0


if __name__ == "__main__":
    sys.exit(main())
