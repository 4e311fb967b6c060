"""Builds the HIP engine in-tree: csrc/*.hip -> optlab/native/liboptlab_mi.so for gfx950.

    python option-pricing-lab_amd/build.py [--jobs N]

Each translation unit is compiled separately (in parallel) and linked once. Packed FP32 instructions are disabled
(-packed-fp32-ops off): plain float code only, the compiler may not fuse lanes into v_pk_* forms. FP32 division and
square root are the hardware's (1 ulp) and FP32 denormals are flushed, as the reference's nvcc flags did."""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "optlab" / "native" / "liboptlab_mi.so"
BUILD = HERE / "_build"
ARCH = os.environ.get("OPTLAB_HIP_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-fvisibility=hidden",
    # FP32 arithmetic as the reference GPU engine compiled it (-prec-div=false -prec-sqrt=false -ftz=true,
    # packages/native/cuda/CMakeLists.txt:50): division by v_rcp_f32, sqrt by v_sqrt_f32, FP32 denormals flushed.
    # FP64 keeps IEEE division, square root and denormals.
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-fgpu-flush-denormals-to-zero",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm install)")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [hipcc(), *FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int = 0) -> Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(len(sources), os.cpu_count() or 1, 8)
    with cf.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(_compile, sources))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-lhiprtc", "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, OUT)
    return OUT


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--jobs", type=int, default=0)
    args = ap.parse_args(argv)
    print(build(args.jobs))
    return 0


if __name__ == "__main__":
    sys.exit(main())
