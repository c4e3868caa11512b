"""Build recipe for the oracle's C part (``oracle/cpu_cic.c`` -> ``oracle/libpm_cpu_cic.so``),
the OpenMP CIC stages of the timed CPU baseline.  Test/measurement infrastructure only.

    python -m oracle.build_cpu
"""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "libpm_cpu_cic.so"


def build(verbose: bool = False) -> Path:
    cmd = ["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", str(HERE / "cpu_cic.c"), "-o",
           str(LIB.with_suffix(".so.tmp")), "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"gcc failed for cpu_cic.c:\n{res.stdout}\n{res.stderr}")
    LIB.with_suffix(".so.tmp").replace(LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
