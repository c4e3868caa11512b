"""Build the HIP library ``libpmpp_amd.so`` for gfx950 with hipcc (in-tree).

Each translation unit under ``csrc/`` is compiled in parallel, then linked into one
shared library next to this file.  Usage: ``python -m pmpp_amd.build [-v]``.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB_NAME = "libpmpp_amd.so"
SOURCES = ["pm_cic.hip", "pm_sort.hip", "pm_fft.hip", "pm_gravity.hip", "pm_nbody.hip", "pm_analysis.hip", "pm_conv.hip",
           "pm_correction.hip", "pm_ic.hip"]

ARCH = os.environ.get("PMPP_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed-FP32 instructions stay off: ordinary float code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def lib_path() -> Path:
    return HERE / LIB_NAME


def _compile(src: str, build_dir: Path, verbose: bool) -> Path:
    obj = build_dir / (Path(src).stem + ".o")
    cmd = [HIPCC, *COMMON_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    build_dir = HERE / "_build"
    build_dir.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda s: _compile(s, build_dir, verbose), SOURCES))
    out = lib_path()
    tmp = out.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
